"""Pins of the oracle's full-IPC mode (psd = 3, beyond the reference; CPU only):
the closest-feature distance against a bounded minimisation over the two features
(scipy), the gradient and the unprojected Hessian against central differences of
the oracle's own energy and gradient, the PSD projection against numpy eigh, the
mollifier against its closed form, and the body-space assembly against numpy."""
import numpy as np
import pytest
from scipy.optimize import minimize

from helpers import full_ipc_config, params, rel_err, rest_of
from oracle import oracle as O

D2HAT, KAPPA = 1e-6, 1e7
CASES = [(0, "face"), (0, "edge"), (0, "vertex"), (1, "line"), (1, "pe"), (1, "pp"), (1, "parallel")]
EXPECT_TYPE = {"face": 0, "edge": 1, "vertex": 2, "line": 0, "pe": 1, "pp": 2, "parallel": 0}


def _min_dist2(kind, x):
    """Squared distance between the two features by bounded minimisation."""
    if kind == 0:
        p, a, b, c = x

        def f(z):
            u, v = z
            q = a + u * (b - a) + v * (c - a)
            return np.sum((p - q) ** 2)

        cons = [{"type": "ineq", "fun": lambda z: 1 - z[0] - z[1]}]
        best = min((minimize(f, z0, bounds=[(0, 1), (0, 1)], constraints=cons, method="SLSQP",
                             options=dict(ftol=1e-18, maxiter=500)) for z0 in ([0.3, 0.3], [0.1, 0.8], [0.8, 0.1])),
                   key=lambda r: r.fun)
        return best.fun
    a0, a1, b0, b1 = x
    v = b1 - b0

    def f_s(s):  # distance from a(s) to segment b: convex in s
        q = a0 + s * (a1 - a0)
        t = min(max((q - b0) @ v / (v @ v), 0.0), 1.0)
        return np.sum((q - b0 - t * v) ** 2)

    lo, hi = 0.0, 1.0
    for _ in range(200):  # ternary search on the convex f_s
        m1, m2 = lo + (hi - lo) / 3, hi - (hi - lo) / 3
        if f_s(m1) <= f_s(m2):
            hi = m2
        else:
            lo = m1
    return min(f_s(0.5 * (lo + hi)), f_s(0.0), f_s(1.0))


@pytest.mark.parametrize("kind,feature", CASES)
def test_closest_feature_distance_matches_minimisation(kind, feature):
    rng = np.random.default_rng(100 + 7 * kind + len(feature))
    for _ in range(12):
        dist = rng.uniform(0.1, 0.9) * 1e-3
        x = full_ipc_config(rng, kind, feature, dist)
        act, E, g, H, f = O.full_local(kind, x, x, D2HAT, KAPPA, project=False)
        assert act and f[0] == EXPECT_TYPE[feature], (feature, f)
        s_min = _min_dist2(kind, x)
        b = O.barrier(s_min, D2HAT, KAPPA)
        if kind == 1:  # undo the mollifier with rest = current
            u, v = x[1] - x[0], x[3] - x[2]
            eps, c = 1e-3 * (u @ u) * (v @ v), np.sum(np.cross(u, v) ** 2)
            b *= (2 - c / eps) * c / eps if c < eps else 1.0
        assert abs(E - b) <= 1e-6 * abs(b), (feature, E, b)
        assert abs(dist ** 2 - s_min) <= 1e-6 * dist ** 2


@pytest.mark.parametrize("kind,feature", CASES)
def test_gradient_and_hessian_by_central_differences(kind, feature):
    rng = np.random.default_rng(200 + 7 * kind + len(feature))
    for _ in range(4):
        dist = rng.uniform(0.2, 0.8) * 1e-3
        x = full_ipc_config(rng, kind, feature, dist)
        r = rest_of(x, rng, 0.05)
        act, E, g, H, f = O.full_local(kind, x, r, D2HAT, KAPPA, project=False)
        assert act
        h = 1e-10
        fd_g = np.zeros(12)
        fd_H = np.zeros((12, 12))
        for i in range(12):
            xp, xm = x.copy().reshape(-1), x.copy().reshape(-1)
            xp[i] += h
            xm[i] -= h
            fd_g[i] = (O.full_energy(kind, xp.reshape(4, 3), r, D2HAT, KAPPA) -
                       O.full_energy(kind, xm.reshape(4, 3), r, D2HAT, KAPPA)) / (2 * h)
            gp = O.full_local(kind, xp.reshape(4, 3), r, D2HAT, KAPPA, project=False)[2]
            gm = O.full_local(kind, xm.reshape(4, 3), r, D2HAT, KAPPA, project=False)[2]
            fd_H[:, i] = (gp - gm) / (2 * h)
        assert rel_err(g, fd_g) <= 1e-5, (feature, rel_err(g, fd_g))
        assert rel_err(H, fd_H) <= 1e-4, (feature, rel_err(H, fd_H))
        assert rel_err(H, H.T) <= 1e-12


def test_mollifier_closed_form_on_parallel_edges():
    rng = np.random.default_rng(5)
    x = full_ipc_config(rng, 1, "parallel", 5e-4)
    r = x.copy()
    u, v = x[1] - x[0], x[3] - x[2]
    eps, c = 1e-3 * (u @ u) * (v @ v), np.sum(np.cross(u, v) ** 2)
    assert c < eps  # the mollifier is active in this configuration
    act, E, *_ = O.full_local(1, x, r, D2HAT, KAPPA, project=False)
    s = _min_dist2(1, x)
    assert abs(E - O.barrier(s, D2HAT, KAPPA) * (2 - c / eps) * (c / eps)) <= 1e-6 * abs(E)


@pytest.mark.parametrize("kind,feature", CASES)
def test_projection_matches_eigh(kind, feature):
    rng = np.random.default_rng(300 + 7 * kind + len(feature))
    x = full_ipc_config(rng, kind, feature, 4e-4)
    r = rest_of(x, rng, 0.05)
    H = O.full_local(kind, x, r, D2HAT, KAPPA, project=False)[3]
    Hp = O.full_local(kind, x, r, D2HAT, KAPPA, project=True)[3]
    lam, V = np.linalg.eigh(0.5 * (H + H.T))
    ref = (V * np.maximum(lam, 0)) @ V.T
    assert rel_err(Hp, ref) <= 1e-9
    assert np.linalg.eigvalsh(0.5 * (Hp + Hp.T)).min() >= -1e-9 * np.abs(lam).max()


@pytest.mark.parametrize("kind,feature", [(0, "edge"), (1, "pp"), (1, "parallel")])
def test_body_space_assembly(kind, feature):
    rng = np.random.default_rng(400 + kind)
    x = full_ipc_config(rng, kind, feature, 3e-4)
    r = rest_of(x, rng, 0.3)
    prm = O.params(params(psd=3))
    act, s, g24, Hi, Hj, Hij = O.pair_gh(kind, prm, x, r)
    _, E, g12, H12, _ = O.full_local(kind, x, r, prm.d2hat, prm.kappa, project=True)
    nI = 1 if kind == 0 else 2
    J = np.zeros((12, 24))
    for t in range(4):
        off = 0 if t < nI else 12
        Jt = np.concatenate([[1.0], r[t]])
        for a in range(4):
            J[3 * t:3 * t + 3, off + 3 * a:off + 3 * a + 3] = Jt[a] * np.eye(3)
    H24 = J.T @ H12 @ J
    assert act
    assert rel_err(g24, J.T @ g12) <= 1e-12
    assert rel_err(np.block([[Hi, Hij], [Hij.T, Hj]]), H24) <= 1e-12
