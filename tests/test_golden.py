"""The oracle reproduces the committed golden fixtures (tests/golden/make_golden.py),
and the fixed reference input behaves as the closed forms require."""
import os

import numpy as np

from abd_hip import scenes
from abd_hip.const_params import ContactParams
from oracle import oracle as O
from refmath import line_line_distance, point_plane_distance

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_fixed_pair_fixture():
    f = np.load(os.path.join(G, "fixed_pair.npz"))
    x = f["x"]
    prm = O.params(ContactParams(dhat=float(f["dhat"])))
    for kind, tag in ((0, "pt"), (1, "ee")):
        geo = O.pair_geometry(kind, x)
        act, s, *_ = O.pair_gh(kind, prm, x, x)
        assert not act and not geo["valid"] and not bool(f[f"{tag}_valid"])
        np.testing.assert_allclose(np.concatenate([geo["e0p"], geo["e1p"], geo["e2p"]]), f[f"{tag}_c"], rtol=1e-13,
                                   atol=1e-15)
        np.testing.assert_allclose(geo["Q"], f[f"{tag}_Q"], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(geo["lam"], f[f"{tag}_lam"], rtol=1e-12, atol=1e-14)
        for X, R, g_ref, H_ref in zip(f[f"{tag}_act_x"], f[f"{tag}_act_r"], f[f"{tag}_act_g24"], f[f"{tag}_act_H24"]):
            act, s, g24, Hi, Hj, Hij = O.pair_gh(kind, prm, X, R)
            assert act
            np.testing.assert_allclose(g24, g_ref, rtol=0, atol=1e-12 * np.abs(g_ref).max())
            H = np.block([[Hi, Hij], [Hij.T, Hj]])
            np.testing.assert_allclose(H, H_ref, rtol=0, atol=1e-12 * np.abs(H_ref).max())
    # closed forms on the fixed input: |e2p| is the point-plane / line-line distance
    assert abs(np.linalg.norm(f["pt_c"][6:]) - point_plane_distance(*x)) <= 1e-12
    assert abs(np.linalg.norm(f["ee_c"][6:]) - line_line_distance(*x)) <= 1e-12


def test_cubic_kat_fixture():
    f = np.load(os.path.join(G, "cubic_kat.npz"))
    for c, known, n, r in zip(f["coef"], f["known"], f["n"], f["roots"]):
        n2, r2 = O.cubic_roots(c, -5.0, 5.0)
        assert n2 == n
        np.testing.assert_array_equal(r2, r)
        if n == 3:
            np.testing.assert_allclose(np.sort(r2), known, atol=1e-8)


def test_small_frames_fixture():
    f = np.load(os.path.join(G, "small_frames.npz"))
    spec = scenes.config_small()
    sim = O.OracleSim(spec, ContactParams(ground=spec.ground, solver_tol=1e-12))
    for k in range(f["A"].shape[0]):
        assert sim.step() == f["iters"][k]
        np.testing.assert_allclose(sim.A, f["A"][k], rtol=0, atol=1e-12)
        np.testing.assert_allclose(sim.p, f["p"][k], rtol=0, atol=1e-12)
