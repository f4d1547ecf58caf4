"""Implicit-function-theorem adjoint of the simulation (differentiation
configuration: 16 icosphere-2 bodies dropped onto a floor, 50 frames, loss
sum_b |com_b - target_b|^2, gradient with respect to the 16 x 3 initial
velocities), against central differences (h = 1e-6) of the GPU forward — all
96 perturbed runs advanced as one batch of independent scenes — and against
central differences of the oracle's forward along a random direction."""
import time

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 50
H = 1e-6


def _params():
    from abd_hip.const_params import ContactParams
    # Exact Hessian; Newton and the linear solves run to convergence, since the adjoint
    # is exact for the minimiser of every frame.  Jacobi CG (the Hessian is SPD here: only
    # vertex-ground contacts) at 1e-12 reaches final states within 1e-13 of those of a
    # Newton tolerance of 1e-20 (measured on the oracle), far below what central
    # differences with h = 1e-6 resolve.  Unpreconditioned BiCGSTAB at a tolerance its
    # conditioning (~1e9) cannot reach stagnates instead, and its wandering iterates can
    # derail a landing frame.
    return ContactParams(ground=0.0, psd=2, solver=1, solver_tol=1e-12, solver_maxiter=20000, tol=1e-8,
                         max_iter=100)


def _loss_np(A, p, c, t):
    com = p + np.einsum("bij,bi->bj", A, c)
    return float(((com - t) ** 2).sum())


@pytest.fixture(scope="module")
def setup():
    from abd_hip import scenes
    from abd_hip.adjoint import Rollout, RolloutFunction, centre_of_mass, rest_centroids
    spec = scenes.config_diff()
    targets = scenes.diff_targets(spec)
    cp = _params()
    ro = Rollout(spec, cp)
    c = rest_centroids(spec)
    t = torch.as_tensor(targets, device="cuda")
    v0 = torch.as_tensor(spec.pdot, dtype=torch.float64, device="cuda").clone().requires_grad_(True)
    A, p = RolloutFunction.apply(v0, ro, STEPS)
    L = ((centre_of_mass(A, p, c) - t) ** 2).sum()
    L.backward()
    return dict(spec=spec, targets=targets, cp=cp, ro=ro, grad=v0.grad.detach().cpu().numpy(), L=L.item(),
                c=c.cpu().numpy())


def test_adjoint_matches_batched_central_differences(setup):
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    spec, cp = setup["spec"], setup["cp"]
    nb = spec.n_bodies
    runs = []
    for i in range(3 * nb):
        for sgn in (1.0, -1.0):
            sp = scenes.subset_scene(spec, range(nb))
            sp.pdot = spec.pdot.copy()
            sp.pdot[i // 3, i % 3] += sgn * H
            runs.append(sp)
    batch = scenes.concat_scenes(runs)
    sim = AffineBodySimulator(batch, cp, "cuda", native=True)
    for f in range(STEPS):
        sim.step(f)
    A = sim.states.A0.cpu().numpy()
    p = sim.states.p0.cpu().numpy()
    fd = np.zeros(3 * nb)
    for i in range(3 * nb):
        Lp = _loss_np(A[(2 * i) * nb:(2 * i + 1) * nb], p[(2 * i) * nb:(2 * i + 1) * nb], setup["c"], setup["targets"])
        Lm = _loss_np(A[(2 * i + 1) * nb:(2 * i + 2) * nb], p[(2 * i + 1) * nb:(2 * i + 2) * nb], setup["c"],
                      setup["targets"])
        fd[i] = (Lp - Lm) / (2 * H)
    g = setup["grad"].reshape(-1)
    err = np.abs(g - fd).max() / np.abs(fd).max()
    print(f"adjoint vs batched central differences: max rel err {err:.2e}, |g|max {np.abs(g).max():.3e}")
    assert err <= 1e-5, err
    # the bodies reached the floor: the gradient passes through contact
    assert sim.stats["gh_active"] + sim.stats["gh_vg"] > 0


def test_adjoint_matches_oracle_differences(setup):
    """Directional derivative along a seeded unit direction: GPU adjoint vs central
    differences of the oracle's forward (same exact-Hessian mode)."""
    import copy
    spec, cp = setup["spec"], setup["cp"]
    rng = np.random.default_rng(5)
    u = rng.normal(size=spec.pdot.shape)
    u /= np.linalg.norm(u)
    Ls = []
    for sgn in (1.0, -1.0):
        sp = copy.deepcopy(spec)
        sp.pdot = spec.pdot + sgn * H * u
        ref = O.OracleSim(sp, cp)
        for _ in range(STEPS):
            ref.step()
        Ls.append(_loss_np(ref.A, ref.p, setup["c"], setup["targets"]))
    fd = (Ls[0] - Ls[1]) / (2 * H)
    g = setup["grad"]
    gu = float((g * u).sum())
    print(f"directional: adjoint {gu:.9e} oracle FD {fd:.9e}")
    assert abs(gu - fd) <= 1e-5 * np.linalg.norm(g), (gu, fd)


def test_adjoint_costs_at_most_three_forward_passes(setup):
    from abd_hip.adjoint import RolloutFunction, centre_of_mass
    ro = setup["ro"]
    spec = setup["spec"]
    c = torch.as_tensor(setup["c"], device="cuda")
    t = torch.as_tensor(setup["targets"], device="cuda")
    v0 = torch.as_tensor(spec.pdot, dtype=torch.float64, device="cuda").clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A, p = RolloutFunction.apply(v0, ro, STEPS)
    L = ((centre_of_mass(A, p, c) - t) ** 2).sum()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = t1 - t0, t2 - t1
    print(f"forward {1e3 * fwd:.1f} ms ({sum(ro.newton_iters)} Newton iterations), backward {1e3 * bwd:.1f} ms "
          f"({sum(ro.solver_iters)} solver iterations), ratio {bwd / fwd:.2f}")
    assert np.array_equal(v0.grad.cpu().numpy(), setup["grad"])
    assert bwd <= 3.0 * fwd
