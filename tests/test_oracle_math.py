"""Pins of the oracle's per-contact math (CPU only): barrier derivatives by
finite differences, contact-space maps by closed-form distances, validity by
constructed inside/outside points (test_distance.py:17-81), the analytic
eigensystem by numpy eigh of the reference's block formula for d2Psi
(psd/hl.py:20-70), the gradient by finite differences of the barrier energy in
affine coordinates, and the 12x12 assembly by a numpy restatement of the
reference's dense host path (ipcsk/pt_ipc.py:194-246, ipcsk/ee_ipc.py:146-211)."""
import numpy as np
import pytest

from helpers import near_contact_ee, near_contact_pt, params, rel_err, rest_of
from oracle import oracle as O
from refmath import affine_map, barycentric, d2psi, line_line_distance, point_plane_distance, signed_distance

D2HAT, KAPPA = 1e-6, 1e7


# ---------------------------------------------------------------- barrier
def test_barrier_closed_form_and_zero_at_dhat():
    # b(s) = -kappa (s/s^ - 1)^2 ln(s/s^): at s = s^/e it is kappa (1 - 1/e)^2
    s = D2HAT / np.e
    assert abs(O.barrier(s, D2HAT, KAPPA) - KAPPA * (1 - 1 / np.e) ** 2) <= 1e-12 * KAPPA
    for f in (O.barrier, O.barrier_d1, O.barrier_d2):
        assert f(D2HAT, D2HAT, KAPPA) == 0.0
        assert f(2 * D2HAT, D2HAT, KAPPA) == 0.0
    assert O.barrier(0.0, D2HAT, KAPPA) == np.inf


@pytest.mark.parametrize("s", [1e-9, 3e-8, 2e-7, 5e-7, 9e-7])
def test_barrier_derivatives_by_finite_differences(s):
    h = 1e-6 * s
    fd1 = (O.barrier(s + h, D2HAT, KAPPA) - O.barrier(s - h, D2HAT, KAPPA)) / (2 * h)
    fd2 = (O.barrier_d1(s + h, D2HAT, KAPPA) - O.barrier_d1(s - h, D2HAT, KAPPA)) / (2 * h)
    assert abs(O.barrier_d1(s, D2HAT, KAPPA) - fd1) <= 1e-6 * abs(fd1)
    assert abs(O.barrier_d2(s, D2HAT, KAPPA) - fd2) <= 1e-6 * abs(fd2)
    assert O.barrier_d1(s, D2HAT, KAPPA) < 0 < O.barrier_d2(s, D2HAT, KAPPA)


# ----------------------------------------------------------- distance maps
def test_pt_map_distance_orthogonality_and_barycentrics():
    rng = np.random.default_rng(0)
    for _ in range(200):
        x = rng.uniform(-1, 1, size=(4, 3))
        G = O.pair_geometry(0, x)
        e0, e1p, e2p = G["e0p"], G["e1p"], G["e2p"]
        p, t0, t1, t2 = x
        # psd/vf.py maps: e0 = x1 - x2, e1 = x3 - x2, e2 = x0 - x2
        assert abs(np.linalg.norm(e2p) - point_plane_distance(p, t0, t1, t2)) <= 1e-12
        assert abs(e1p @ e0) <= 1e-12 * np.linalg.norm(e0) ** 2 * np.linalg.norm(e1p) + 1e-14
        assert abs(e2p @ e0) <= 1e-11 and abs(e2p @ (t2 - t1)) <= 1e-11
        assert abs(G["l"] ** 2 - e2p @ e2p) <= 1e-12  # psd/hl.py:254-265
        assert abs(G["l"] - signed_distance(e0, e1p, p - t1)) <= 1e-12
        # beta, gamma: foot of p = t1 + beta (t0 - t1) + gamma (t2 - t1)
        uv = barycentric(p, t1, t0, t2)
        np.testing.assert_allclose([G["beta"], G["gamma"]], uv, atol=1e-9)


def test_ee_map_distance():
    rng = np.random.default_rng(1)
    for _ in range(200):
        x = rng.uniform(-1, 1, size=(4, 3))
        G = O.pair_geometry(1, x)
        assert abs(np.linalg.norm(G["e2p"]) - line_line_distance(*x)) <= 1e-11
        assert abs(abs(G["l"]) - line_line_distance(*x)) <= 1e-11
        # closest points: x0 + beta e0 and x2 - gamma e1 coincide up to e2p
        a = x[0] + G["beta"] * (x[1] - x[0])
        b = x[2] - G["gamma"] * (x[3] - x[2])
        np.testing.assert_allclose(b - a, G["e2p"], atol=1e-10)


def _triangle_case(rng, inside):
    # test_distance.py:17-54 init_positive / init_negative with a per-case RNG
    e0, e1 = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
    b = rng.uniform(0.01, 0.99)
    g = rng.uniform(0.01, 0.99) * (1 - b)
    if not inside:
        b, g = 1.0 - g, 1.0 - b
    n = np.cross(e0, e1)
    t1 = np.zeros(3)
    return np.stack([n * rng.uniform() + t1 + b * e0 + g * e1, t1, e0 + t1, e1 + t1])


def test_pt_inside_invariant():
    """test_distance.py:68-81: points projecting outside the triangle are invalid,
    inside valid (1000 cases each, independent seeds per case)."""
    rng = np.random.default_rng(114514)
    for inside in (False, True):
        for _ in range(1000):
            x = _triangle_case(rng, inside)
            # the reference's vertex order is (p, t0, t1, t2) with e0 = t1 - t2 ... use its fetch order
            assert O.pair_geometry(0, x[[0, 2, 1, 3]])["valid"] == inside


def test_parallel_edges_invalid():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1e-4, 0], [1, 1e-4, 0]], np.float64)
    assert not O.pair_geometry(1, x)["valid"]  # singular 2x2 -> zero inverse -> beta = 0 (R4)
    x = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0, 0], [0.5, 1, 1e-4]], np.float64)
    assert not O.pair_geometry(0, x[[3, 0, 1, 2]])["valid"]  # degenerate triangle


# ---------------------------------------------------------- eigensystem
@pytest.mark.parametrize("kind", [0, 1])
def test_eigensystem_reproduces_d2psi(kind):
    """Sum_k lam_k q_k q_k^T / |q_k|^2 equals the reference's block formula of
    d2Psi (EE, unclamped) and eigh's PSD projection of it (PT, clamped)."""
    rng = np.random.default_rng(2 + kind)
    for _ in range(100):
        d = rng.uniform(-1e-3, 1e-3)
        x = near_contact_pt(rng, d) if kind == 0 else near_contact_ee(rng, d)
        G = O.pair_geometry(kind, x)
        Q, lam = G["Q"], G["lam"]
        H = sum(lam[k] * np.outer(Q[k], Q[k]) / (Q[k] @ Q[k]) for k in range(5))
        D = d2psi(G["e0p"], G["e1p"], G["e2p"])
        if kind == 0:
            w, V = np.linalg.eigh(D)
            D = V @ np.diag(np.maximum(w, 0)) @ V.T
        assert rel_err(H, D) <= 1e-10
        # eigenvector property for every pair
        Dfull = d2psi(G["e0p"], G["e1p"], G["e2p"])
        for k in range(5):
            if kind == 1 or k == 4:
                np.testing.assert_allclose(Dfull @ Q[k], lam[k] * Q[k], atol=1e-9 * np.abs(Dfull).max())


# ------------------------------------------------------------ gradient
@pytest.mark.parametrize("kind", [0, 1])
def test_gradient_is_derivative_of_barrier_energy(kind):
    """g24 = d b(s(x(q))) / dq in the 24 affine coordinates (central differences)."""
    rng = np.random.default_rng(4 + kind)
    cp = params()
    prm = O.params(cp)
    nI = 1 if kind == 0 else 2
    checked = 0
    for _ in range(40):
        d = rng.uniform(2e-4, 9e-4)
        xt = near_contact_pt(rng, d) if kind == 0 else near_contact_ee(rng, d)
        # affine state per body (p, rows of A near identity); rest positions chosen so
        # that x(q) reproduces the contact configuration: x = p + r @ A
        q = np.zeros(24)
        r4 = np.zeros((4, 3))
        for half, verts in ((0, range(nI)), (12, range(nI, 4))):
            A = np.eye(3) + 0.2 * rng.uniform(-1, 1, size=(3, 3))
            p = rng.uniform(-1, 1, 3)
            q[half:half + 3] = p
            q[half + 3:half + 12] = A.reshape(-1)
            for v in verts:
                r4[v] = np.linalg.solve(A.T, xt[v] - p)
        x = affine_map(q, r4, nI)
        act, s, g24, *_ = O.pair_gh(kind, prm, x, r4)
        if not act:
            continue

        def E(qq):
            xx = affine_map(qq, r4, nI)
            G = O.pair_geometry(kind, xx)
            s2 = G["l"] ** 2 if kind == 0 else G["e2p"] @ G["e2p"]
            return O.barrier(s2, cp.d2hat, cp.kappa)
        h = 1e-8
        fd = np.array([(E(q + h * np.eye(24)[i]) - E(q - h * np.eye(24)[i])) / (2 * h) for i in range(24)])
        assert rel_err(g24, fd) <= 1e-5, rel_err(g24, fd)
        checked += 1
    assert checked >= 20


# ----------------------------------------------- dense reference assembly
def _ref_dense(kind, G, r4, B1, B2):
    """ipcsk/pt_ipc.py:194-246 / ipcsk/ee_ipc.py:146-211 in numpy: extract_Q,
    QLQinv, dcTHldc, extract_g, JTg, JTH12J with np.kron."""
    Qn = G["Q"].reshape(15, 3)
    Qm = Qn.reshape(5, 9).T
    QTQ = Qm.T @ Qm
    Qinv = np.diag(1.0 / np.diag(QTQ)) @ Qm.T
    Hl = Qm @ np.diag(G["lam"]) @ Qinv
    K = np.kron(G["D"], np.eye(3))
    H12 = K.T @ Hl @ K
    g12 = np.kron(G["D"].T, np.eye(3)) @ Qn[12:15].reshape(-1)
    Hipc = H12 * B1 + np.outer(g12, g12) * B2
    nI = 1 if kind == 0 else 2
    JI = np.hstack([np.ones((nI, 1)), r4[:nI]])
    JJ = np.hstack([np.ones((4 - nI, 1)), r4[nI:]])
    gi = np.kron(JI.T, np.eye(3)) @ g12[:3 * nI]
    gj = np.kron(JJ.T, np.eye(3)) @ g12[3 * nI:]
    KI, KJ = np.kron(JI, np.eye(3)), np.kron(JJ, np.eye(3))
    a = 3 * nI
    Hi = KI.T @ Hipc[:a, :a] @ KI
    Hj = KJ.T @ Hipc[a:, a:] @ KJ
    Hij = KI.T @ Hipc[:a, a:] @ KJ
    return np.concatenate([gi, gj]) * B1, Hi, Hj, Hij


@pytest.mark.parametrize("kind", [0, 1])
def test_pair_assembly_matches_numpy_restatement(kind):
    rng = np.random.default_rng(6 + kind)
    cp = params()
    prm = O.params(cp)
    for _ in range(60):
        d = rng.uniform(-9e-4, 9e-4)
        x = near_contact_pt(rng, d) if kind == 0 else near_contact_ee(rng, d)
        r4 = rest_of(x, rng)
        act, s, g24, Hi, Hj, Hij = O.pair_gh(kind, prm, x, r4)
        assert act
        G = O.pair_geometry(kind, x)
        B1, B2 = O.barrier_d1(s, cp.d2hat, cp.kappa), O.barrier_d2(s, cp.d2hat, cp.kappa)
        g_ref, Hi_r, Hj_r, Hij_r = _ref_dense(kind, G, r4, B1, B2)
        assert rel_err(g24, g_ref) <= 1e-12
        H = np.block([[Hi, Hij], [Hij.T, Hj]])
        Hr = np.block([[Hi_r, Hij_r], [Hij_r.T, Hj_r]])
        assert rel_err(H, Hr) <= 1e-12


@pytest.mark.parametrize("kind", [0, 1])
def test_psd_mode_is_projection_of_contact_space_hessian(kind):
    """psd=1 (beyond the reference): the 24x24 body-pair block is J^T D^T P+(Hc) D J with
    Hc = b' d2Psi + b'' q4 q4^T and P+ the PSD projection by numpy eigh."""
    rng = np.random.default_rng(8 + kind)
    cp = params(psd=1)
    prm = O.params(cp)
    for _ in range(40):
        d = rng.uniform(-9e-4, 9e-4)
        x = near_contact_pt(rng, d) if kind == 0 else near_contact_ee(rng, d)
        r4 = rest_of(x, rng)
        act, s, g24, Hi, Hj, Hij = O.pair_gh(kind, prm, x, r4)
        G = O.pair_geometry(kind, x)
        B1, B2 = O.barrier_d1(s, cp.d2hat, cp.kappa), O.barrier_d2(s, cp.d2hat, cp.kappa)
        Dpsi = d2psi(G["e0p"], G["e1p"], G["e2p"])
        q4 = G["Q"][4]
        Hc = B1 * Dpsi + B2 * np.outer(q4, q4)
        w, V = np.linalg.eigh(0.5 * (Hc + Hc.T))
        Hp = V @ np.diag(np.maximum(w, 0)) @ V.T
        nI = 1 if kind == 0 else 2
        K = np.kron(G["D"], np.eye(3))
        Jb = np.zeros((12, 24))
        for v in range(4):
            row = np.concatenate([[1.0], r4[v]])
            off = 0 if v < nI else 12
            for a in range(4):
                Jb[3 * v:3 * v + 3, off + 3 * a:off + 3 * a + 3] = row[a] * np.eye(3)
        Href = Jb.T @ K.T @ Hp @ K @ Jb
        H = np.block([[Hi, Hij], [Hij.T, Hj]])
        assert rel_err(H, Href) <= 1e-9
        assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() >= -1e-9 * np.abs(H).max()


@pytest.mark.parametrize("kind", [0, 1])
def test_exact_mode_hessian_is_derivative_of_gradient(kind):
    """psd=2 (beyond the reference): the 24x24 body-pair block is the exact Hessian of
    b(s(x(q))) — central differences of the exact gradient g24 in the 24 affine
    coordinates — and the gradient is the reference's."""
    rng = np.random.default_rng(12 + kind)
    cp = params(psd=2)
    prm, prm0 = O.params(cp), O.params(params())
    nI = 1 if kind == 0 else 2
    checked = 0
    for _ in range(40):
        d = rng.uniform(2e-4, 9e-4)
        xt = near_contact_pt(rng, d) if kind == 0 else near_contact_ee(rng, d)
        q = np.zeros(24)
        r4 = np.zeros((4, 3))
        for half, verts in ((0, range(nI)), (12, range(nI, 4))):
            A = np.eye(3) + 0.2 * rng.uniform(-1, 1, size=(3, 3))
            p = rng.uniform(-1, 1, 3)
            q[half:half + 3] = p
            q[half + 3:half + 12] = A.reshape(-1)
            for v in verts:
                r4[v] = np.linalg.solve(A.T, xt[v] - p)
        x = affine_map(q, r4, nI)
        act, s, g24, Hi, Hj, Hij = O.pair_gh(kind, prm, x, r4)
        if not act:
            continue
        _, _, g_ref, *_ = O.pair_gh(kind, prm0, x, r4)
        assert rel_err(g24, g_ref) <= 1e-9
        H = np.block([[Hi, Hij], [Hij.T, Hj]])
        assert rel_err(H, H.T) <= 1e-12
        h = 1e-7
        fd = np.zeros((24, 24))
        for i in range(24):
            e = np.eye(24)[i] * h
            gp = O.pair_gh(kind, prm, affine_map(q + e, r4, nI), r4)[2]
            gm = O.pair_gh(kind, prm, affine_map(q - e, r4, nI), r4)[2]
            fd[:, i] = (gp - gm) / (2 * h)
        assert rel_err(H, fd) <= 1e-5, rel_err(H, fd)
        checked += 1
    assert checked >= 20


def test_exact_mode_vg_hessian_includes_first_derivative_term():
    """psd=2: the vertex-ground block is (2 b' + 4 d^2 b'') theta theta^T e_y e_y^T (the
    exact second derivative of b(d^2)); the reference keeps only the b'' term (D11)."""
    from abd_hip import scenes
    cp = params(psd=2)
    spec = scenes.config_small()
    sc = O.OracleScene(spec)
    x = O.project(sc, spec.A, spec.p)
    x = x.copy()
    d = 4e-4
    x[0, 1] = cp.ground + d
    vg = np.array([[0, 0]], np.int32)
    g0, B0 = O.contact_gh(sc, O.params(params()), x, None, None, vg, 0)[:2]
    g2, B2 = O.contact_gh(sc, O.params(cp), x, None, None, vg, 0)[:2]
    assert np.array_equal(g0, g2)
    b1 = O.barrier_d1(d * d, cp.d2hat, cp.kappa)
    b2 = O.barrier_d2(d * d, cp.d2hat, cp.kappa)
    th = np.concatenate([[1.0], spec.x0[0]])
    # (ii, jj) sub-block at 16 b + ii + 4 jj, entry (1, 1)
    for ii in range(4):
        for jj in range(4):
            got = B2[ii + 4 * jj][1, 1] - B0[ii + 4 * jj][1, 1]
            assert abs(got - th[ii] * th[jj] * 2.0 * b1) <= 1e-9 * abs(th[ii] * th[jj] * 2.0 * b1) + 1e-300
