"""Full-IPC mode (psd = 3, beyond the reference: closest feature pairs, edge-edge
mollifier, per-contact 12x12 PSD projection) on the GPU against the oracle's
independent restatement (brute-force feature choice, forward-mode derivatives,
classical Jacobi): per contact 1e-9 relative, and Newton frames of the native
driver to 1e-7."""
import numpy as np
import pytest
import torch

from helpers import full_ipc_config, params, rel_err, rest_of
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(0, "face"), (0, "edge"), (0, "vertex"), (1, "line"), (1, "pe"), (1, "pp"), (1, "parallel")]


@pytest.mark.parametrize("kind", [0, 1])
def test_pair_gh_full_ipc_matches_oracle(kind):
    import abd_hip
    from abd_hip.ipc import pair_gh
    abd_hip.load()
    rng = np.random.default_rng(30 + kind)
    cp = params(psd=3)
    X, R = [], []
    for k, feat in CASES:
        if k != kind:
            continue
        for _ in range(24):
            x4 = full_ipc_config(rng, kind, feat, rng.uniform(0.05, 0.95) * 1e-3)
            X.append(x4)
            R.append(rest_of(x4, rng, 0.05))
    for _ in range(8):  # far: inactive
        x4 = full_ipc_config(rng, kind, "face" if kind == 0 else "line", 0.1)
        X.append(x4)
        R.append(rest_of(x4, rng, 0.05))
    X, R = np.array(X), np.array(R)
    act, g24, H24 = pair_gh(kind, cp, torch.as_tensor(X, device="cuda"), torch.as_tensor(R, device="cuda"))
    torch.cuda.synchronize()
    act, g24, H24 = act.cpu().numpy(), g24.cpu().numpy(), H24.cpu().numpy()
    prm = O.params(cp)
    n_act = 0
    for i in range(len(X)):
        a, s, g, Hi, Hj, Hij = O.pair_gh(kind, prm, X[i], R[i])
        assert bool(act[i]) == a, i
        if not a:
            continue
        n_act += 1
        Href = np.block([[Hi, Hij], [Hij.T, Hj]])
        assert rel_err(g24[i], g) <= 1e-9, (i, rel_err(g24[i], g))
        assert rel_err(H24[i], Href) <= 1e-9, (i, rel_err(H24[i], Href))
    assert n_act == len(X) - 8


def test_medium_subset_full_ipc_frames():
    """Newton frames of the native driver in full-IPC mode against the oracle's driver."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_medium(seed=0, n_side=2, subdiv=2)
    spec.p[:] = spec.p * 0.93
    cp = ContactParams(ground=spec.ground, psd=3, solver=2, solver_tol=1e-10, solver_maxiter=20000)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    ref = O.OracleSim(spec, cp)
    for f in range(10):
        it = sim.step(f)
        it_ref = ref.step()
        torch.cuda.synchronize()
        A, p = sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()
        err = max(np.abs(A - ref.A).max(), np.abs(p - ref.p).max())
        assert err <= 1e-7, (f, err, it, it_ref)
    assert sim.stats["gh_active"] > 0
