"""Pins of the oracle's culling, CCD, inertia and solver stages (CPU only):
boxes against a numpy restatement of the FP32 cast rules (culling2.py:33-280),
candidate sets against brute-force numpy overlap, the cubic root solver against
the reference's known-root construction (cubic_roots.py:225-269) and sign-change
brute force, TOI against closed forms, inertia derivatives by finite differences,
and the Krylov solvers against numpy.linalg.solve."""
import numpy as np
import pytest

from abd_hip import scenes
from helpers import contact_scene, params, rel_err
from oracle import oracle as O


def _boxes_numpy(spec, x, kind, dil):
    """culling2.py R3 rules in numpy (float32 casts are round-to-nearest)."""
    if kind == 0:
        c = x.astype(np.float32)
        return c - np.float32(dil), c + np.float32(dil)
    topo, offs = (spec.edges, spec.eoff) if kind == 1 else (spec.tris, spec.foff)
    body = np.repeat(np.arange(spec.n_bodies), np.diff(offs))
    gv = topo + spec.voff[body][:, None]
    pts = x[gv]
    return (pts.min(1) - dil).astype(np.float32), (pts.max(1) + dil).astype(np.float32)


def _overlap_pairs(lo1, hi1, lo2, hi2):
    ok = np.all((lo1[:, None, :] <= hi2[None, :, :]) & (hi1[:, None, :] >= lo2[None, :, :]), axis=2)
    return np.argwhere(ok)


@pytest.mark.parametrize("seed", [0, 1])
def test_boxes_and_candidate_sets_brute_force(seed):
    spec = contact_scene(seed=seed)
    rng = np.random.default_rng(seed)
    spec.p[:] += 1e-4 * rng.normal(size=spec.p.shape)
    sc = O.OracleScene(spec)
    x = O.project(sc, spec.A, spec.p)
    # projection: x = A^T x0 + p per body
    for b in range(spec.n_bodies):
        v0, v1 = spec.voff[b], spec.voff[b + 1]
        np.testing.assert_allclose(x[v0:v1], spec.x0[v0:v1] @ spec.A[b] + spec.p[b], rtol=0, atol=1e-14)
    dhat = 1e-3
    boxes = {}
    for kind, dil in ((0, dhat), (1, dhat / 2), (2, 0.0)):
        lo, hi = O.prim_boxes(sc, kind, x, dil)
        nlo, nhi = _boxes_numpy(spec, x, kind, dil)
        np.testing.assert_array_equal(lo, nlo)
        np.testing.assert_array_equal(hi, nhi)
        boxes[kind] = (lo, hi)
    blo, bhi = O.body_boxes(sc, x, dhat / 2)
    for b in range(spec.n_bodies):
        xb = x[spec.voff[b]:spec.voff[b + 1]]
        np.testing.assert_array_equal(blo[b], (xb.min(0) - dhat / 2).astype(np.float32))
        np.testing.assert_array_equal(bhi[b], (xb.max(0) + dhat / 2).astype(np.float32))
    ij = O.cull_bodies(sc, blo, bhi)
    ref_ij = [(i, j) for i, j in _overlap_pairs(blo, bhi, blo, bhi) if i < j]
    assert [tuple(r) for r in ij] == sorted(ref_ij)
    # EE and PT sets
    ee = O.cull_prims(False, ij, sc.eoff, *boxes[1], sc.eoff, *boxes[1])
    pt = O.cull_prims(True, ij, sc.voff, *boxes[0], sc.foff, *boxes[2])
    exp_ee, exp_pt = [], []
    for ii, (I, J) in enumerate(ij):
        e0, e1 = spec.eoff[I], spec.eoff[I + 1]
        f0, f1 = spec.eoff[J], spec.eoff[J + 1]
        for a, b in _overlap_pairs(boxes[1][0][e0:e1], boxes[1][1][e0:e1], boxes[1][0][f0:f1], boxes[1][1][f0:f1]):
            exp_ee.append((I, J, ii, a, b))
        for (P, T) in ((I, J), (J, I)):
            v0, v1 = spec.voff[P], spec.voff[P + 1]
            t0, t1 = spec.foff[T], spec.foff[T + 1]
            for a, b in _overlap_pairs(boxes[0][0][v0:v1], boxes[0][1][v0:v1], boxes[2][0][t0:t1],
                                       boxes[2][1][t0:t1]):
                exp_pt.append((P, T, ii, a, b))
    assert [tuple(r) for r in ee] == exp_ee
    assert [tuple(r) for r in pt] == exp_pt
    assert len(ee) > 0 and len(pt) > 0
    vg = O.cull_vg(sc, blo, *boxes[0], dhat, spec.ground)
    exp_vg = [(b, v - spec.voff[b]) for b in range(spec.n_bodies) if float(blo[b, 1]) < 0.5 * dhat + spec.ground
              for v in range(spec.voff[b], spec.voff[b + 1])
              if boxes[0][0][v, 1] <= np.float32(spec.ground) and np.all(boxes[0][0][v] <= 1e3)
              and np.all(boxes[0][1][v] >= -1e3)]
    assert [tuple(r) for r in vg] == exp_vg and len(vg) > 0


def test_trajectory_box_union():
    spec = contact_scene()
    sc = O.OracleScene(spec)
    rng = np.random.default_rng(3)
    x0 = O.project(sc, spec.A, spec.p)
    x1 = x0 + 0.01 * rng.normal(size=x0.shape)
    for kind, dil in ((0, 1e-3), (1, 5e-4), (2, 0.0)):
        lo, hi = O.prim_boxes(sc, kind, x0, dil)
        tlo, thi = O.prim_boxes_traj(sc, kind, x1, lo, hi)
        l1, h1 = _boxes_numpy(spec, x1, kind, 0.0)
        np.testing.assert_array_equal(tlo, np.minimum(lo, l1))
        np.testing.assert_array_equal(thi, np.maximum(hi, h1))


# -------------------------------------------------------------------- CCD
def test_cubic_known_roots():
    """cubic_roots.py:225-246 construction: three roots in [0, bound), scaled by a0,
    solved on [-5, 5] (cubic_roots.py:201)."""
    rng = np.random.default_rng(0)
    for _ in range(500):
        r = np.sort(rng.uniform(0, 5.0, 3))
        if np.min(np.diff(r)) < 1e-3:
            continue
        a0 = rng.uniform(0.1, 5.0)
        coef = np.array([-r[0] * r[1] * r[2], r[0] * r[1] + r[0] * r[2] + r[1] * r[2], -r.sum(), 1.0]) * a0
        n, roots = O.cubic_roots(coef, -5.0, 5.0)
        assert n == 3
        np.testing.assert_allclose(np.sort(roots), r, atol=1e-8)


def test_cubic_roots_against_sign_changes():
    rng = np.random.default_rng(1)
    t = np.linspace(0.0, 1.0, 20001)
    for _ in range(500):
        coef = rng.normal(size=4)
        n, roots = O.cubic_roots(coef, 0.0, 1.0)
        y = coef[0] + t * (coef[1] + t * (coef[2] + t * coef[3]))
        changes = int(np.sum(y[1:] * y[:-1] < 0))
        assert n >= min(changes, 1)
        for k in range(n):
            assert -1e-12 <= roots[k] <= 1 + 1e-12
            yr = coef[0] + roots[k] * (coef[1] + roots[k] * (coef[2] + roots[k] * coef[3]))
            assert abs(yr) <= 1e-8 * np.abs(coef).sum()
        if n:
            # the first root returned is a genuine sign change nearest to the solver's bracket
            assert changes >= 1


def test_point_triangle_toi_closed_form():
    rng = np.random.default_rng(2)
    for _ in range(200):
        t = rng.uniform(-1, 1, size=(3, 3))
        n = np.cross(t[1] - t[0], t[2] - t[0])
        if np.linalg.norm(n) < 0.2:
            continue
        n /= np.linalg.norm(n)
        d0 = rng.uniform(0.01, 0.1)
        v = rng.uniform(0.02, 0.3)
        p0 = t[0] + 0.3 * (t[1] - t[0]) + 0.3 * (t[2] - t[0]) + d0 * n
        P0 = np.stack([p0, t[0], t[1], t[2]])
        P1 = P0.copy()
        P1[0] = p0 - v * n
        toi = O.collision_time(0, P0, P1)
        exp = d0 / v if d0 / v <= 1 else 1.0
        assert abs(toi - exp) <= 1e-9, (toi, exp)


def test_edge_edge_toi_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a0, d0, d1 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        n = np.cross(d0, d1)
        if np.linalg.norm(n) < 0.3:
            continue
        n /= np.linalg.norm(n)
        gap, v = rng.uniform(0.01, 0.1), rng.uniform(0.02, 0.3)
        c = a0 + 0.4 * d0 + gap * n
        b0 = c - 0.6 * d1
        P0 = np.stack([a0, a0 + d0, b0, b0 + d1])
        P1 = P0.copy()
        P1[2:] -= v * n
        toi = O.collision_time(1, P0, P1)
        exp = gap / v if gap / v <= 1 else 1.0
        assert abs(toi - exp) <= 1e-9, (toi, exp)


def test_vertex_ground_toi_closed_form():
    spec = scenes.config_small()
    sc = O.OracleScene(spec)
    cp = params(ground=spec.ground)
    x0 = O.project(sc, spec.A, spec.p)
    x1 = x0.copy()
    x1[:, 1] -= 0.01
    vg = np.array([[0, v] for v in range(8) if spec.x0[v, 1] < 0], np.int32)
    d0 = x0[vg[0, 1], 1] - spec.ground
    d1 = x1[vg[0, 1], 1] - spec.ground
    assert abs(O.toi(sc, O.params(cp), x0, x1, None, None, vg) - d0 / (d0 - d1)) <= 1e-15


# ---------------------------------------------------------------- inertia
def test_inertia_derivatives_by_finite_differences():
    spec = scenes.config_medium(seed=2, n_side=2, subdiv=1)
    prm = O.params(params())
    rng = np.random.default_rng(4)
    nb = spec.n_bodies
    A = spec.A + 0.05 * rng.normal(size=spec.A.shape)
    p = spec.p + 0.05 * rng.normal(size=spec.p.shape)
    Adot = 0.5 * rng.normal(size=spec.A.shape)
    args = (spec.A, spec.p, Adot, spec.pdot, spec.mass, spec.I0)

    def pack(A, p):
        q = np.zeros((nb, 4, 3))
        q[:, 0] = p
        q[:, 1:] = A
        return q.reshape(-1)

    def unpack(q):
        q = q.reshape(nb, 4, 3)
        return q[:, 1:].copy(), q[:, 0].copy()

    E = lambda q: O.inertia_energy(prm, *unpack(q), *args)
    G = lambda q: O.inertia_gradient(prm, *unpack(q), *args).reshape(-1)
    q = pack(A, p)
    g = G(q)
    h = 1e-6
    idx = rng.choice(q.size, 40, replace=False)
    for i in idx:
        e = np.zeros(q.size)
        e[i] = h
        fd = (E(q + e) - E(q - e)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd))
    H = O.inertia_hessian(prm, A, np.zeros((16 * nb, 3, 3)), spec.mass, spec.I0)
    rows, cols = O.set_triplets(nb, np.zeros((0, 2), np.int32))
    Hd = np.zeros((12 * nb, 12 * nb))
    for t in range(H.shape[0]):
        Hd[3 * rows[t]:3 * rows[t] + 3, 3 * cols[t]:3 * cols[t] + 3] += H[t]
    for i in idx[:15]:
        e = np.zeros(q.size)
        e[i] = h
        fd = (G(q + e) - G(q - e)) / (2 * h)
        assert rel_err(Hd[:, i], fd) <= 1e-6


# ----------------------------------------------------------------- solvers
def _system(rng, nb=12, npair=20):
    pairs = sorted({tuple(sorted(rng.choice(nb, 2, replace=False))) for _ in range(npair)})
    ij = np.array(pairs, np.int32)
    nij = len(ij)
    blocks = np.zeros(((nb + 2 * nij) * 16, 3, 3))
    dense = np.zeros((12 * nb, 12 * nb))

    def put(slot, M):
        for ii in range(4):
            for jj in range(4):
                blocks[16 * slot + ii + 4 * jj] = M[3 * ii:3 * ii + 3, 3 * jj:3 * jj + 3]
    for k, (i, j) in enumerate(ij):
        C = 0.3 * rng.normal(size=(12, 12))
        put(nb + k, C)
        put(nb + nij + k, C.T)
        dense[12 * i:12 * i + 12, 12 * j:12 * j + 12] += C
        dense[12 * j:12 * j + 12, 12 * i:12 * i + 12] += C.T
    for b in range(nb):
        Q = rng.normal(size=(12, 12))
        D = Q @ Q.T + 12.0 * (1 + nij) * np.eye(12)
        put(b, D)
        dense[12 * b:12 * b + 12, 12 * b:12 * b + 12] += D
    return ij, blocks, dense


def test_triplets_assemble_the_dense_matrix():
    rng = np.random.default_rng(5)
    nb = 12
    ij, blocks, dense = _system(rng, nb)
    rows, cols = O.set_triplets(nb, ij)
    Hd = np.zeros_like(dense)
    for t in range(blocks.shape[0]):
        Hd[3 * rows[t]:3 * rows[t] + 3, 3 * cols[t]:3 * cols[t] + 3] += blocks[t]
    np.testing.assert_array_equal(Hd, dense)


@pytest.mark.parametrize("which", ["bicgstab", "cg", "cg_block"])
def test_solvers_against_numpy(which):
    rng = np.random.default_rng(6)
    nb = 12
    ij, blocks, dense = _system(rng, nb)
    b = rng.normal(size=(4 * nb, 3))
    x0 = np.zeros_like(b)
    fn = {"bicgstab": O.bicgstab, "cg": O.bsr_cg, "cg_block": O.bsr_cg_block}[which]
    x, it, res = fn(nb, ij, blocks, b, x0, 1e-12)
    xs = np.linalg.solve(dense, b.reshape(-1)).reshape(-1, 3)
    assert rel_err(x, xs) <= 1e-9
    assert 0 < it <= 12 * nb
    # relative tolerance semantics of the BiCGSTAB call (abd.py:244): |r| <= tol |b|
    x, it, res = O.bicgstab(nb, ij, blocks, b, x0, 1e-3)
    r = b.reshape(-1) - dense @ x.reshape(-1)
    assert np.linalg.norm(r) <= 1e-3 * np.linalg.norm(b)


def _chain_system(rng, nb, stiff):
    """Bodies in a chain (i, i+1), each link a stiff PSD coupling: ill-conditioned like a
    compressed stack.  Returns (ij, blocks, dense)."""
    ij = np.array([(i, i + 1) for i in range(nb - 1)], np.int32)
    nij = len(ij)
    blocks = np.zeros(((nb + 2 * nij) * 16, 3, 3))
    dense = np.zeros((12 * nb, 12 * nb))

    def put(slot, M):
        for ii in range(4):
            for jj in range(4):
                blocks[16 * slot + ii + 4 * jj] += M[3 * ii:3 * ii + 3, 3 * jj:3 * jj + 3]
    Dg = [np.eye(12) * 10.0 for _ in range(nb)]
    for k, (i, j) in enumerate(ij):
        v = rng.normal(size=(12, 3))
        K = stiff * (v @ v.T)  # the spring of link k between the two bodies' dofs
        Dg[i] += K
        Dg[j] += K
        put(nb + k, -K)
        put(nb + nij + k, -K.T)
        dense[12 * i:12 * i + 12, 12 * j:12 * j + 12] -= K
        dense[12 * j:12 * j + 12, 12 * i:12 * i + 12] -= K.T
    for b in range(nb):
        put(b, Dg[b])
        dense[12 * b:12 * b + 12, 12 * b:12 * b + 12] += Dg[b]
    return ij, blocks, dense


@pytest.mark.parametrize("nb,npair", [(1, 0), (2, 1), (12, 20), (40, 120), (64, 200)])
def test_ldlt_against_numpy(nb, npair):
    """The direct block LDL^T (beyond the reference) against numpy's dense solve, on random
    couplings (fill-in included) of up to the 64 bodies it takes."""
    rng = np.random.default_rng(nb)
    ij, blocks, dense = _system(rng, nb, npair) if nb > 1 else _system(rng, 2, 1)
    if nb == 1:
        ij, nb = np.zeros((0, 2), np.int32), 1
        blocks = blocks[:16].copy()
        dense = dense[:12, :12].copy()
    b = rng.normal(size=(4 * nb, 3))
    x, st = O.bsr_ldlt(nb, ij, blocks, b, np.zeros_like(b))
    assert st == 1
    xs = np.linalg.solve(dense, b.reshape(-1)).reshape(-1, 3)
    assert rel_err(x, xs) <= 1e-10


def test_ldlt_on_a_stiff_chain():
    """A chain of stiff links: the block-Jacobi CG needs about a hundred iterations, the
    factorization is exact to rounding (a chain has no fill-in)."""
    rng = np.random.default_rng(3)
    nb = 48
    ij, blocks, dense = _chain_system(rng, nb, 1e6)
    b = rng.normal(size=(4 * nb, 3))
    x, st = O.bsr_ldlt(nb, ij, blocks, b, np.zeros_like(b))
    assert st == 1
    xs = np.linalg.solve(dense, b.reshape(-1))
    r = b.reshape(-1) - dense @ x.reshape(-1)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(dense, 2) * np.linalg.norm(xs)
    assert rel_err(x.reshape(-1), xs) <= 1e-6
    _, it, _ = O.bsr_cg_block(nb, ij, blocks, b, np.zeros_like(b), 1e-10)
    assert it >= 90


def test_ldlt_declines_what_it_cannot_factor():
    """A non-positive pivot (indefinite diagonal block) and a system above 64 bodies give
    status -2 with x untouched."""
    rng = np.random.default_rng(4)
    nb = 6
    ij, blocks, dense = _system(rng, nb, 8)
    blocks[16 * 2:16 * 3] *= -1.0  # body 2's diagonal block negative definite
    b = rng.normal(size=(4 * nb, 3))
    x0 = rng.normal(size=b.shape)
    x, st = O.bsr_ldlt(nb, ij, blocks, b, x0)
    assert st == -2 and np.array_equal(x, x0)
    nb = 65
    ij, blocks, dense = _chain_system(rng, nb, 1.0)
    b = rng.normal(size=(4 * nb, 3))
    x, st = O.bsr_ldlt(nb, ij, blocks, b, np.zeros_like(b))
    assert st == -2 and not x.any()


def test_cg_guard_when_the_residual_vanishes_between_checks():
    """Defect D16 (bsr_utils.py:70 divides rz_old / pAp unguarded): with a diagonal matrix the
    Jacobi preconditioner is its exact inverse, the residual is exactly 0 after the first
    iteration, and the reference's next step would be 0/0.  The guard skips the step: the
    result is the exact solution, finite, stopping at the first check."""
    rng = np.random.default_rng(12)
    nb = 3
    ij = np.array([(0, 1)], np.int32)
    blocks = np.zeros(((nb + 2) * 16, 3, 3))
    d = np.exp2(rng.integers(0, 4, size=(nb, 12))).astype(float)  # powers of two: exact quotients
    for b in range(nb):
        for ii in range(4):
            blocks[16 * b + ii + 4 * ii] = np.diag(d[b, 3 * ii:3 * ii + 3])
    rhs = np.exp2(rng.integers(-3, 3, size=(4 * nb, 3))).astype(float)
    x, it, res = O.bsr_cg(nb, ij, blocks, rhs, np.zeros_like(rhs), 1e-12)
    assert np.isfinite(x).all() and res == 0.0 and it == 9
    np.testing.assert_array_equal(x.reshape(nb, 12), rhs.reshape(nb, 12) / d)


def test_block_jacobi_cg_is_preconditioned_cg():
    """The block-Jacobi CG (beyond the reference) equals a numpy restatement of PCG with
    M = blockdiag(12x12 diagonal blocks), iterate for iterate, and takes fewer iterations
    than the diagonal preconditioner when bodies are strongly coupled inside their block."""
    rng = np.random.default_rng(8)
    nb = 10
    ij, blocks, dense = _system(rng, nb)
    n = 12 * nb
    # make the 12x12 diagonal blocks strongly non-diagonal (what contacts do to a body)
    for q in range(nb):
        C = rng.normal(size=(12, 12))
        C = 40.0 * (C @ C.T)
        for ii in range(4):
            for jj in range(4):
                blocks[16 * q + ii + 4 * jj] += C[3 * ii:3 * ii + 3, 3 * jj:3 * jj + 3]
        dense[12 * q:12 * q + 12, 12 * q:12 * q + 12] += C
    b = rng.normal(size=(4 * nb, 3))
    x0 = np.zeros_like(b)
    Minv = np.zeros((n, n))
    for q in range(nb):
        Minv[12 * q:12 * q + 12, 12 * q:12 * q + 12] = np.linalg.inv(dense[12 * q:12 * q + 12, 12 * q:12 * q + 12])
    # numpy PCG with the same stopping rule (tol^2 n on r.z, every 10 iterations)
    tol, bb = 1e-10, b.reshape(-1)
    x = np.zeros(n); r = bb - dense @ x; z = Minv @ r; p = z.copy(); rz = r @ z
    it_ref = 0
    for i in range(n):
        Ap = dense @ p; a = rz / (p @ Ap); x += a * p; r -= a * Ap; z = Minv @ r; rz_new = r @ z
        if (i + 1) % 10 == 0 and rz_new <= tol * tol * n:
            it_ref = i
            break
        p = z + (rz_new / rz) * p; rz = rz_new
    xo, it, _ = O.bsr_cg_block(nb, ij, blocks, b, x0, tol)
    assert it == it_ref
    assert rel_err(xo.reshape(-1), x) <= 1e-9
    _, it_diag, _ = O.bsr_cg(nb, ij, blocks, b, x0, tol)
    assert it < it_diag
