"""The separation checks of tests/geometry.py against brute force (dense sampling
of the primitives) and constructed cases, and on the oracle's two-cube run."""
import numpy as np

import geometry as G


def test_point_triangle_distance_matches_sampling():
    rng = np.random.default_rng(0)
    u, v = np.meshgrid(np.linspace(0, 1, 201), np.linspace(0, 1, 201))
    m = u + v <= 1
    u, v = u[m], v[m]
    for _ in range(40):
        t = rng.uniform(-1, 1, (3, 3))
        p = rng.uniform(-1.5, 1.5, 3)
        pts = t[0] + u[:, None] * (t[1] - t[0]) + v[:, None] * (t[2] - t[0])
        brute = np.linalg.norm(pts - p, axis=1).min()
        d = G.point_triangle_dist(p, t[0], t[1], t[2])
        assert d <= brute + 1e-12
        assert brute - d < 2e-2 * max(np.linalg.norm(t[1] - t[0]), np.linalg.norm(t[2] - t[0]))


def test_segment_distance_matches_sampling():
    rng = np.random.default_rng(1)
    s = np.linspace(0, 1, 1001)
    for _ in range(40):
        a0, a1, b0, b1 = rng.uniform(-1, 1, (4, 3))
        A = a0 + s[:, None] * (a1 - a0)
        B = b0 + s[:, None] * (b1 - b0)
        brute = np.linalg.norm(A[:, None] - B[None], axis=-1).min()
        d = G.segment_segment_dist(a0, a1, b0, b1)
        assert d <= brute + 1e-12 and brute - d < 5e-3
    # parallel segments: the distance of the lines where they overlap
    assert abs(G.segment_segment_dist(np.zeros(3), np.array([1.0, 0, 0]), np.array([0.5, 0.2, 0]),
                                      np.array([2.0, 0.2, 0])) - 0.2) < 1e-15


def test_segment_triangle_crossing():
    t = [np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])]
    assert G.segment_crosses_triangle(np.array([0.2, 0.2, -1]), np.array([0.2, 0.2, 1]), *t)
    assert not G.segment_crosses_triangle(np.array([0.2, 0.2, 0.1]), np.array([0.2, 0.2, 1]), *t)
    assert not G.segment_crosses_triangle(np.array([0.8, 0.8, -1]), np.array([0.8, 0.8, 1]), *t)


def test_separation_of_resting_and_overlapping_cubes():
    from abd_hip import scenes
    from abd_hip.const_params import ContactParams
    from oracle import oracle as O
    spec = scenes.config_small()
    cp = ContactParams(ground=spec.ground)
    sim = O.OracleSim(spec, cp)
    for _ in range(60):
        sim.step()
    x = O.project(O.OracleScene(spec), sim.A, sim.p)
    r = G.separation(spec, x, spec.ground)
    assert r["pairs"] == 1 and r["crossings"] == 0
    assert 0 < min(r["min_pt"], r["min_ee"]) < cp.dhat          # resting on B within dhat
    assert 0 < r["min_vg"] < cp.dhat                             # B resting on the ground
    # the same cubes pushed 0.1 into each other: edges cross faces
    p = sim.p.copy()
    p[1, 1] -= 0.1
    x = O.project(O.OracleScene(spec), sim.A, p)
    assert G.separation(spec, x, spec.ground)["crossings"] > 0
