"""Newton driver (abd.py:212-276) on the GPU against the oracle's driver: body
states (A, p) agree to 1e-7 after every frame.  The linear solve runs to a tight
tolerance (1e-12 relative) in these runs so that both sides follow the same
Newton path; the per-stage tests cover the 1e-4 production setting."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _run(spec, cp, frames, native=False):
    from abd_hip.abd import AffineBodySimulator
    sim = AffineBodySimulator(spec, cp, "cuda", native=native)
    ref = O.OracleSim(spec, cp)
    worst = 0.0
    for f in range(frames):
        it = sim.step(f)
        it_ref = ref.step()
        torch.cuda.synchronize()
        A, p = sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()
        err = max(np.abs(A - ref.A).max(), np.abs(p - ref.p).max())
        worst = max(worst, err)
        assert err <= 1e-7, (f, err, it, it_ref)
    return sim, ref, worst


@pytest.mark.parametrize("native", [False, True])
def test_small_two_cubes_100_frames(native):
    from abd_hip import scenes
    from abd_hip.const_params import ContactParams
    spec = scenes.config_small()
    cp = ContactParams(ground=spec.ground, solver_tol=1e-12)
    sim, ref, worst = _run(spec, cp, 100, native)
    # A came to rest on B: gap of about dhat, no interpenetration
    assert 0.0 < sim.states.p[1, 1].item() - sim.states.p[0, 1].item() - 1.0 < 2e-3
    assert sim.stats["newton_iters"] == ref.stats[0]
    # every Newton iteration's collision set is the culling's (PT, EE, VG sizes summed)
    assert [sim.stats[k] for k in ("gh_pt", "gh_ee", "gh_vg")] == [int(v) for v in ref.stats[1:4]]


@pytest.mark.parametrize("native", [False, True])
def test_medium_subset_frames(native):
    from abd_hip import scenes
    from abd_hip.const_params import ContactParams
    spec = scenes.config_medium(seed=0, n_side=2, subdiv=2)
    spec.p[:] = spec.p * 0.93
    cp = ContactParams(ground=spec.ground, solver_tol=1e-12)
    sim, ref, worst = _run(spec, cp, 8, native)
    assert sim.stats["gh_pt"] + sim.stats["gh_ee"] > 0
    # the same Newton iterations over the same collision sets (PT, EE, VG sizes summed)
    assert sim.stats["newton_iters"] == ref.stats[0]
    assert [sim.stats[k] for k in ("gh_pt", "gh_ee", "gh_vg")] == [int(v) for v in ref.stats[1:4]]


def test_small_frames_match_golden_fixture():
    import os
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_frames.npz"))
    spec = scenes.config_small()
    sim = AffineBodySimulator(spec, ContactParams(ground=spec.ground, solver_tol=1e-12), "cuda")
    for k in range(f["A"].shape[0]):
        sim.step(k)
        torch.cuda.synchronize()
        assert np.abs(sim.states.A.cpu().numpy() - f["A"][k]).max() <= 1e-7
        assert np.abs(sim.states.p.cpu().numpy() - f["p"][k]).max() <= 1e-7


def test_medium_subset_psd_cg_frames():
    """The stable benchmark mode (contact-space PSD projection + Jacobi CG) against
    the oracle in the same mode."""
    from abd_hip import scenes
    from abd_hip.const_params import ContactParams
    spec = scenes.config_medium(seed=0, n_side=2, subdiv=2)
    spec.p[:] = spec.p * 0.93
    # CG runs to convergence (maxiter well above n) so that both sides reach the same solution
    cp = ContactParams(ground=spec.ground, psd=1, solver=1, solver_tol=1e-10, solver_maxiter=20000)
    sim, ref, worst = _run(spec, cp, 10, native=True)
    assert sim.stats["gh_active"] > 0


def test_medium_subset_block_jacobi_cg_frames():
    """The stable mode with the block-Jacobi CG (beyond the reference) against the oracle in
    the same mode."""
    from abd_hip import scenes
    from abd_hip.const_params import ContactParams
    spec = scenes.config_medium(seed=0, n_side=2, subdiv=2)
    spec.p[:] = spec.p * 0.93
    cp = ContactParams(ground=spec.ground, psd=1, solver=2, solver_tol=1e-10, solver_maxiter=20000)
    sim, ref, worst = _run(spec, cp, 10, native=True)
    assert sim.stats["gh_active"] > 0


def _pin_cases():
    from helpers import driver_pin_cases
    return driver_pin_cases()


@pytest.mark.parametrize("native", [False, True])
@pytest.mark.parametrize("name", sorted(_pin_cases()))
def test_closed_form_frame(name, native):
    """The driver's control logic against closed forms (no oracle in the loop): free fall,
    the pre-update stop test, max_iter, the as-written Armijo test, halving, alpha_min, and
    the CCD caps against the ground and between two cubes (cases in tests/helpers.py)."""
    from abd_hip.abd import AffineBodySimulator
    spec, cp, dq, p_want, it_want, tol = _pin_cases()[name]
    print(f"pin case {name} native={native}", flush=True)
    sim = AffineBodySimulator(spec, cp, "cuda", native=native)
    if dq is not None:
        sim.dq.copy_(torch.as_tensor(dq, device="cuda"))
    it = sim.step(0)
    torch.cuda.synchronize()
    assert it == it_want
    np.testing.assert_allclose(sim.states.p.cpu().numpy(), p_want, rtol=0, atol=tol)
    if name.startswith("free_fall"):
        np.testing.assert_array_equal(sim.states.A.cpu().numpy(), np.eye(3)[None])


def test_two_simulators_stepped_from_two_threads():
    """Each simulator owns its scratch: two native simulators stepped at the same time
    from two threads, each on its own stream, follow the trajectories they follow alone."""
    import threading
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams

    def specs():
        a = scenes.config_medium(seed=0, n_side=2, subdiv=2)
        a.p[:] = a.p * 0.93
        return [a, scenes.config_batched(n_scenes=6, seed0=11)]

    frames = 12

    def run(sim, out, stream):
        with torch.cuda.stream(stream):
            for f in range(frames):
                sim.step(f)
            stream.synchronize()
            out.append((sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()))

    alone = []
    for spec in specs():
        cp = ContactParams(ground=spec.ground, solver=3, psd=1)
        out = []
        sim = AffineBodySimulator(spec, cp, "cuda", native=True)
        torch.cuda.synchronize()
        run(sim, out, torch.cuda.Stream())
        alone.append(out[0])
    sims = [AffineBodySimulator(spec, ContactParams(ground=spec.ground, solver=3, psd=1), "cuda", native=True)
            for spec in specs()]
    torch.cuda.synchronize()
    outs = [[], []]
    threads = [threading.Thread(target=run, args=(sims[k], outs[k], torch.cuda.Stream())) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for k in range(2):
        assert outs[k], f"simulator {k} did not finish"
        np.testing.assert_array_equal(outs[k][0][0], alone[k][0])
        np.testing.assert_array_equal(outs[k][0][1], alone[k][1])
