"""The direct block LDL^T solve (ABD_SOLVER_LDLT, beyond the reference) through
libabd_hip.so: against the oracle's restatement (same elimination order and sums) and
numpy's dense solve; its block-Jacobi-CG fallback (systems above 64 bodies, a pivot that is not
positive); one workgroup per scene in a batch against each system solved alone; and the
Newton driver in this mode against the oracle's driver in the same mode."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda"


def T(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV, dtype=dtype)


def N(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def lib():
    import abd_hip
    return abd_hip.load()


def _system(nb, ij, rng, spd=True):
    """Random SPD (or, spd=False, indefinite diagonal blocks) block system on pairs ij:
    (blocks in the slot layout, dense matrix)."""
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
        D = Q @ Q.T + 12.0 * np.eye(12) * (1 + len(ij))
        if not spd and b == nb // 2:
            D = -D
        put(b, D)
        dense[12 * b:12 * b + 12, 12 * b:12 * b + 12] += D
    return blocks, dense


def _pairs(rng, nb, npair):
    return np.array(sorted({tuple(sorted(rng.choice(nb, 2, replace=False))) for _ in range(npair)}),
                    np.int32).reshape(-1, 2)


def _structure(lib, nb, ij):
    from abd_hip import _lib
    ji = ij[np.lexsort((ij[:, 0], ij[:, 1]))] if len(ij) else ij
    nij = len(ij)
    row_ptr = torch.empty(nb + 1, dtype=torch.int32, device=DEV)
    cols = torch.empty(nb + 2 * nij, dtype=torch.int32, device=DEV)
    slots = torch.empty(nb + 2 * nij, dtype=torch.int32, device=DEV)
    tij, tji = T(ij.reshape(-1, 2)), T(ji.reshape(-1, 2))
    _lib.call("abd_bsr_structure", _lib.stream_handle(), nb, _lib.ptr(tij), _lib.ptr(tji), nij,
              _lib.ptr(row_ptr), _lib.ptr(cols), _lib.ptr(slots))
    return row_ptr, cols, slots


@pytest.mark.parametrize("nb,npair", [(1, 0), (2, 1), (24, 40), (64, 150), (64, 0)])
def test_ldlt_matches_oracle(lib, nb, npair):
    from abd_hip.sparse import BsrMatrix, SolveInfo, bsr_ldlt
    rng = np.random.default_rng(100 + nb + npair)
    ij = _pairs(rng, nb, npair) if npair else np.zeros((0, 2), np.int32)
    blocks, dense = _system(nb, ij, rng)
    b = rng.normal(size=(4 * nb, 3))
    x0 = 0.01 * rng.normal(size=(4 * nb, 3))
    row_ptr, cols, slots = _structure(lib, nb, ij)
    A = BsrMatrix(nb, row_ptr, cols, slots, T(blocks))
    x = T(x0)
    info = SolveInfo(DEV)
    bsr_ldlt(A, x, T(b), 0, 1e-4, 10, info)
    x_ref, st = O.bsr_ldlt(nb, ij, blocks, b, x0)
    assert st == 1 and int(N(info.iters)[0]) == 1
    xs = np.linalg.solve(dense, b.reshape(-1)).reshape(-1, 3)
    assert rel_err(N(x), xs) <= 1e-10
    # same elimination order, same sums, no contraction on either side
    assert rel_err(N(x), x_ref) <= 1e-14
    # resid = |b - H x|^2 of the computed x: rounding level
    assert 0.0 <= float(N(info.resid)[0]) <= 1e-20 * float((b * b).sum())


@pytest.mark.parametrize("case", ["indefinite", "large"])
def test_ldlt_falls_back_to_block_jacobi_cg(lib, case):
    """A system the factorization declines is solved by the block-Jacobi CG exactly as
    method 2 would solve it."""
    from abd_hip.sparse import BsrMatrix, SolveInfo, bsr_cg, bsr_ldlt
    rng = np.random.default_rng(7 if case == "indefinite" else 8)
    nb = 10 if case == "indefinite" else 70
    ij = _pairs(rng, nb, 2 * nb)
    blocks, dense = _system(nb, ij, rng, spd=case != "indefinite")
    if case == "indefinite":
        assert O.bsr_ldlt(nb, ij, blocks, np.ones((4 * nb, 3)), np.zeros((4 * nb, 3)))[1] == -2
    b = rng.normal(size=(4 * nb, 3))
    x0 = 0.01 * rng.normal(size=(4 * nb, 3))
    row_ptr, cols, slots = _structure(lib, nb, ij)
    A = BsrMatrix(nb, row_ptr, cols, slots, T(blocks))
    x1, x2 = T(x0), T(x0)
    i1, i2 = SolveInfo(DEV), SolveInfo(DEV)
    bsr_ldlt(A, x1, T(b), 500, 1e-8, 10, i1)
    bsr_cg(A, x2, T(b), 500, 1e-8, 10, i2, block_precond=True)
    assert np.array_equal(N(x1), N(x2))
    assert int(N(i1.iters)[0]) == int(N(i2.iters)[0]) > 1


def test_ldlt_batched_equals_each_system_alone(lib):
    """abd_bsr_solve_scenes with method 3: one workgroup per scene (sizes 1..64, one of 70
    bodies through the fallback, one masked out and left untouched) gives each scene the
    solution of its own system (the oracle's, to 1e-14; the fallback's to the CG's 1e-7)."""
    from abd_hip import _lib
    rng = np.random.default_rng(11)
    sizes = [1, 5, 64, 17, 70, 33, 2]
    inactive = 5
    boff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    nb = int(boff[-1])
    pairs = []
    for s, m in enumerate(sizes):
        if m > 1:
            pairs += [tuple(p + boff[s]) for p in _pairs(rng, m, 2 * m).tolist()]
    ij = np.array(sorted(set(pairs)), np.int32).reshape(-1, 2)
    blocks, dense = _system(nb, ij, rng)
    b = rng.normal(size=(4 * nb, 3))
    x0 = 0.01 * rng.normal(size=(4 * nb, 3))
    row_ptr, cols, slots = _structure(lib, nb, ij)
    body_scene = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    tboff, tbs = T(boff), T(body_scene)
    sc = _lib.AbdScene()
    sc.n_bodies = nb
    sc.n_scenes = len(sizes)
    sc.scene_boff = tboff.data_ptr()
    sc.body_scene = tbs.data_ptr()
    active = np.ones(len(sizes), np.uint8)
    active[inactive] = 0
    tact = T(active)
    tb, x = T(blocks), T(x0)
    iters = torch.zeros(len(sizes), dtype=torch.int32, device=DEV)
    resid = torch.zeros(len(sizes), dtype=torch.float64, device=DEV)
    _lib.call("abd_bsr_solve_scenes", _lib.stream_handle(), 3, ctypes.byref(sc), max(sizes), _lib.ptr(row_ptr),
              _lib.ptr(cols), _lib.ptr(slots), _lib.ptr(tb), _lib.ptr(T(b)), _lib.ptr(x), 1e-8, 500, 10,
              _lib.ptr(tact), _lib.ptr(iters), _lib.ptr(resid))
    xb, itb = N(x).reshape(-1, 12), N(iters)
    for s, m in enumerate(sizes):
        r0, r1 = boff[s], boff[s + 1]
        if s == inactive:
            assert np.array_equal(xb[r0:r1], x0.reshape(-1, 12)[r0:r1])
            continue
        # the scene's own system: its pairs renumbered from 0
        sel = (ij[:, 0] >= r0) & (ij[:, 0] < r1)
        ijs = ij[sel] - r0
        slots_s = list(range(r0, r1)) + [nb + k for k in np.nonzero(sel)[0]] + \
            [nb + len(ij) + k for k in np.nonzero(sel)[0]]
        bs = blocks.reshape(-1, 16, 3, 3)[slots_s].reshape(-1, 3, 3)
        ref, st = O.bsr_ldlt(m, ijs, bs, b[4 * r0:4 * r1], x0[4 * r0:4 * r1])
        if m <= 64:
            assert st == 1 and itb[s] == 1
            assert rel_err(xb[r0:r1].reshape(-1, 3), ref) <= 1e-14
        else:
            assert st == -2 and itb[s] > 1
            xc, itc, _ = O.bsr_cg_block(m, ijs, bs, b[4 * r0:4 * r1], x0[4 * r0:4 * r1], 1e-8, 500)
            assert rel_err(xb[r0:r1].reshape(-1, 3), xc) <= 1e-7


def test_medium_subset_ldlt_frames():
    """The stable mode with the direct solve against the oracle's driver in the same mode:
    states agree to 1e-7 after every frame."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_medium(seed=0, n_side=2, subdiv=2)
    spec.p[:] = spec.p * 0.93
    cp = ContactParams(ground=spec.ground, psd=1, solver=3)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    ref = O.OracleSim(spec, cp)
    for f in range(10):
        it = sim.step(f)
        it_ref = ref.step()
        A, p = sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()
        err = max(np.abs(A - ref.A).max(), np.abs(p - ref.p).max())
        assert err <= 1e-7, (f, err, it, it_ref)
    assert sim.stats["gh_active"] > 0
    assert sim.stats["solver_iters"] == sim.stats["newton_iters"]  # one direct solve each


def test_batched_ldlt_bitwise_equals_plain():
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    n = 6
    spec = scenes.config_batched(n_scenes=n, seed0=3, nb_range=(2, 8))
    cp = ContactParams(ground=spec.ground, psd=1, solver=3)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    plain = [AffineBodySimulator(scenes.scene_of(spec, s), cp, "cuda", native=True) for s in range(n)]
    for f in range(8):
        sim.step(f)
        its = sim.scene_iters()
        A_all, p_all = sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()
        for s in range(n):
            it = plain[s].step(f)
            assert its[s] == it, (f, s, its[s], it)
            idx = np.nonzero(spec.body_scene == s)[0]
            assert np.array_equal(A_all[idx], plain[s].states.A.cpu().numpy()), (f, s)
            assert np.array_equal(p_all[idx], plain[s].states.p.cpu().numpy()), (f, s)
