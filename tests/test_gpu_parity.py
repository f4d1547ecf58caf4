"""GPU parity of every stage of the contact path against the CPU oracle, through
libabd_hip.so.  Tolerances: candidate lists, boxes, validity and active flags
bit-exact; energies, gradients and Hessians 1e-9 relative (norm-wise: max |diff|
<= 1e-9 max |ref|); TOI 1e-12 absolute (same arithmetic, same root solver)."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import contact_scene, near_contact_ee, near_contact_pt, oracle_lists, params, rel_err, rest_of
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


def _gpu_scene(spec):
    from abd_hip.affine_body import Meshes
    return Meshes(spec, DEV)


# ------------------------------------------------------------------ per contact
@pytest.mark.parametrize("kind", [0, 1])
def test_pair_gh_matches_oracle(lib, kind):
    from abd_hip.ipc import pair_gh
    rng = np.random.default_rng(10 + kind)
    cp = params(dhat=1e-3)
    xs, rs = [], []
    for i in range(256):
        d = rng.uniform(1e-5, 9.9e-4) * (1 if i % 2 else -1)
        x4 = near_contact_pt(rng, d) if kind == 0 else near_contact_ee(rng, d)
        xs.append(x4)
        rs.append(rest_of(x4, rng))
    # inactive (far) and degenerate (parallel edges / collinear) cases too
    for i in range(16):
        x4 = near_contact_pt(rng, 0.5) if kind == 0 else near_contact_ee(rng, 0.5)
        xs.append(x4)
        rs.append(rest_of(x4, rng))
    par = np.array([[0, 0, 0], [1, 0, 0], [0, 1e-4, 0], [1, 1e-4, 0]], np.float64)
    xs.append(par)
    rs.append(par)
    X, R = np.array(xs), np.array(rs)
    act, g24, H24 = pair_gh(kind, cp, T(X), T(R))
    act, g24, H24 = N(act), N(g24), N(H24)
    prm = O.params(cp)
    n_act = 0
    for i in range(len(X)):
        a_ref, s, g_ref, Hi, Hj, Hij = O.pair_gh(kind, prm, X[i], R[i])
        assert bool(act[i]) == a_ref, i
        if not a_ref:
            assert np.all(H24[i] == 0) and np.all(g24[i] == 0)
            continue
        n_act += 1
        Href = np.block([[Hi, Hij], [Hij.T, Hj]])
        assert rel_err(g24[i], g_ref) <= 1e-9, (i, rel_err(g24[i], g_ref))
        assert rel_err(H24[i], Href) <= 1e-9, (i, rel_err(H24[i], Href))
    assert n_act >= 200


# ------------------------------------------------------------- boxes and culling
def _stage_lists(m, cp):
    from abd_hip.culling import BvhBuilder, cull, cull_vg, ij_list
    bb = BvhBuilder(m)
    dh = 0.5 * cp.dhat
    bodies = bb.build_body_bvh(m, dh, field="xk")
    tri = bb.bulid_triangle_bvh(m.xk, m.triangles, 0.0)
    edg = bb.build_edge_bvh(m.xk, m.edges, dh)
    pts = bb.build_point_bvh(m.xk, cp.dhat)
    pairs = ij_list(bodies, m)
    ee = cull(pairs, edg)
    pt = cull(pairs, pts, tri)
    vg = cull_vg(bodies.lowers, pts, m, cp.dhat, cp.ground)
    return dict(bodies=bodies, tri=tri, edg=edg, pts=pts, pairs=pairs, ee=ee, pt=pt, vg=vg)


@pytest.mark.parametrize("which", [0, 1])
def test_per_body_builder_loops_equal_scene_wide(lib, which):
    """The reference's per-body BvhBuilder loops (abd.py:84-93, :171-184, :192-206), run
    unchanged over Meshes' per-body views, give the scene-wide boxes and candidate lists."""
    from abd_hip.culling import BvhBuilder, cull, cull_vg, ij_list
    spec = _scenes()[which]
    cp = params(ground=spec.ground)
    m = _gpu_scene(spec)
    st = m.states()
    from abd_hip import _lib
    _lib.call("abd_project_vertices", _lib.stream_handle(), m.c, _lib.ptr(st.A), _lib.ptr(st.p), _lib.ptr(m.xk))
    m.x.copy_(m.xk)
    m.xk.add_(torch.as_tensor([0.0, -3e-3, 1e-3], dtype=torch.float64, device=DEV))  # a swept step
    dhat, dh = cp.dhat, 0.5 * cp.dhat
    meshes = list(m)  # the reference's per-body mesh list
    assert len(meshes) == spec.n_bodies
    bb = BvhBuilder()
    bvh_bodies = bb.build_body_bvh(meshes, dh)
    bvh_body_traj = bb.build_body_bvh(meshes, dh)
    tris, edges, points, tris_t, edges_t, points_t = [], [], [], [], [], []
    for b in meshes:
        tris.append(bb.bulid_triangle_bvh(b.x, b.triangles, 0.0))
        edges.append(bb.build_edge_bvh(b.x, b.edges, dh))
        points.append(bb.build_point_bvh(b.x, dhat))
        tris_t.append(bb.bulid_triangle_bvh(b.x, b.triangles, 0.0))
        edges_t.append(bb.build_edge_bvh(b.x, b.edges, dh))
        points_t.append(bb.build_point_bvh(b.x, dhat))
    # collision_set (abd.py:192-206)
    bb.update_body_bvh(meshes, dh, bvh_bodies)
    ij = ij_list(bvh_bodies, m)
    for b, t, e, p in zip(meshes, tris, edges, points):
        bb.update_triangle_bvh(b.xk, b.triangles, 0.0, t)
        bb.update_edge_bvh(b.xk, b.edges, dh, e)
        bb.update_point_bvh(b.xk, dhat, p)
    ee, pt = cull(ij, edges), cull(ij, points, tris)
    vg = cull_vg(bvh_bodies.lowers, points, meshes)
    # the same flows scene-wide
    sw = BvhBuilder(m)
    S_bodies = sw.build_body_bvh(m, dh, field="xk")
    S_tri = sw.bulid_triangle_bvh(m.xk, m.triangles, 0.0)
    S_edg = sw.build_edge_bvh(m.xk, m.edges, dh)
    S_pts = sw.build_point_bvh(m.xk, dhat)
    S_ij = ij_list(S_bodies, m)
    assert torch.equal(bvh_bodies.lowers, S_bodies.lowers) and torch.equal(bvh_bodies.uppers, S_bodies.uppers)
    for lst, S in ((tris, S_tri), (edges, S_edg), (points, S_pts)):
        parent = lst[0].parent
        assert all(x.parent is parent for x in lst) and [x.body for x in lst] == list(range(spec.n_bodies))
        assert torch.equal(parent.lowers, S.lowers) and torch.equal(parent.uppers, S.uppers)
    assert torch.equal(S_ij.ij, ij.ij)
    assert torch.equal(cull(S_ij, S_edg), ee) and torch.equal(cull(S_ij, S_pts, S_tri), pt)
    assert torch.equal(cull_vg(S_bodies.lowers, S_pts, m, cp.dhat, cp.ground), vg)
    # the trajectory set (abd.py:170-184): boxes of x, swept to xk
    bb.body_bvh_to_traj(meshes, bb.build_body_bvh(meshes, dh), bvh_body_traj)
    for i, b in enumerate(meshes):
        bb.update_triangle_bvh(b.x, b.triangles, 0.0, tris[i])
        bb.update_edge_bvh(b.x, b.edges, dh, edges[i])
        bb.update_point_bvh(b.x, dhat, points[i])
        bb.triangle_bvh_to_traj(b.xk, b.triangles, tris[i], tris_t[i])
        bb.edge_bvh_to_traj(b.xk, b.edges, edges[i], edges_t[i])
        bb.point_bvh_to_traj(b.xk, points[i], points_t[i])
    ij_t = ij_list(bvh_body_traj, m)
    ee_t, pt_t = cull(ij_t, edges_t), cull(ij_t, points_t, tris_t)

    S_tri_x = sw.bulid_triangle_bvh(m.x, m.triangles, 0.0)
    S_edg_x = sw.build_edge_bvh(m.x, m.edges, dh)
    S_pts_x = sw.build_point_bvh(m.x, dhat)
    S_tri_t = sw.triangle_bvh_to_traj(m.xk, m.triangles, S_tri_x, sw.bulid_triangle_bvh(m.x, m.triangles, 0.0))
    S_edg_t = sw.edge_bvh_to_traj(m.xk, m.edges, S_edg_x, sw.build_edge_bvh(m.x, m.edges, dh))
    S_pts_t = sw.point_bvh_to_traj(m.xk, S_pts_x, sw.build_point_bvh(m.x, dhat))
    for lst, S in ((tris_t, S_tri_t), (edges_t, S_edg_t), (points_t, S_pts_t)):
        parent = lst[0].parent
        assert all(x.parent is parent for x in lst)
        assert torch.equal(parent.lowers, S.lowers) and torch.equal(parent.uppers, S.uppers)
    S_bt = sw.body_bvh_to_traj(m, sw.build_body_bvh(m, dh), sw.build_body_bvh(m, dh))
    assert torch.equal(S_bt.lowers, bvh_body_traj.lowers) and torch.equal(S_bt.uppers, bvh_body_traj.uppers)
    S_ij_t = ij_list(S_bt, m)
    assert torch.equal(cull(S_ij_t, S_edg_t), ee_t) and torch.equal(cull(S_ij_t, S_pts_t, S_tri_t), pt_t)
    assert pt.shape[0] + ee.shape[0] > 0 and pt_t.shape[0] + ee_t.shape[0] > 0


def _scenes():
    from abd_hip import scenes
    med = scenes.config_medium(seed=3, n_side=2)  # 8 icosphere-3 bodies
    # squeeze the grid so that neighbours overlap
    med.p[:] = med.p * 0.9
    small = scenes.config_small()
    small.p[1, 1] = 1.0 + 4e-4  # A resting on B
    return [contact_scene(), med, small]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_boxes_and_culling_bit_exact(lib, which):
    spec = _scenes()[which]
    cp = params(ground=spec.ground)
    m = _gpu_scene(spec)
    st = m.states()
    from abd_hip import _lib
    _lib.call("abd_project_vertices", _lib.stream_handle(), m.c, _lib.ptr(st.A), _lib.ptr(st.p), _lib.ptr(m.xk))
    sc = O.OracleScene(spec)
    x_ref = O.project(sc, spec.A, spec.p)
    np.testing.assert_array_equal(N(m.xk), x_ref)
    G = _stage_lists(m, cp)
    R = oracle_lists(sc, x_ref, cp.dhat, cp.ground)
    np.testing.assert_array_equal(N(G["bodies"].lowers), R["blo"])
    np.testing.assert_array_equal(N(G["bodies"].uppers), R["bhi"])
    for k, rk in (("tri", "t"), ("edg", "e"), ("pts", "p")):
        np.testing.assert_array_equal(N(G[k].lowers), R[rk + "lo"])
        np.testing.assert_array_equal(N(G[k].uppers), R[rk + "hi"])
    np.testing.assert_array_equal(N(G["pairs"].ij), R["ij"])
    ji_ref = R["ij"][np.lexsort((R["ij"][:, 0], R["ij"][:, 1]))] if len(R["ij"]) else R["ij"]
    np.testing.assert_array_equal(N(G["pairs"].ji), ji_ref)
    for k in ("ee", "pt", "vg"):
        np.testing.assert_array_equal(N(G[k]), R[k])
    assert len(R["ij"]) > 0


def test_cull_overflow_reports_the_required_count(lib):
    """Defect D2 (lists capped at 256 with no bound check, fenwick.py:127-136,165-166): a
    list that does not fit its capacity returns ABD_ERR_OVERFLOW with the count it needs and
    writes nothing past the capacity; a second call at that capacity gives the full list."""
    import ctypes
    from abd_hip import _lib
    spec = _scenes()[1]
    cp = params(ground=spec.ground)
    m = _gpu_scene(spec)
    st = m.states()
    _lib.call("abd_project_vertices", _lib.stream_handle(), m.c, _lib.ptr(st.A), _lib.ptr(st.p), _lib.ptr(m.xk))
    G = _stage_lists(m, cp)
    full = N(G["pt"])
    assert len(full) > 8
    ijt = G["pairs"].ij
    out = torch.full((len(full) + 4, 5), -7, dtype=torch.int32, device=DEV)
    cnt = ctypes.c_int64(0)
    cap = len(full) // 2
    rc = _lib.call("abd_cull_prims", _lib.stream_handle(), m.c, _lib.CULL_PT, _lib.ptr(ijt), int(ijt.shape[0]),
                   _lib.ptr(G["pts"].lowers), _lib.ptr(G["pts"].uppers), _lib.ptr(G["tri"].lowers),
                   _lib.ptr(G["tri"].uppers), _lib.ptr(out), cap, ctypes.byref(cnt),
                   ok=(_lib.ABD_OK, _lib.ABD_ERR_OVERFLOW))
    assert rc == _lib.ABD_ERR_OVERFLOW and cnt.value == len(full)
    assert (N(out)[cap:] == -7).all()  # nothing written past the capacity
    rc = _lib.call("abd_cull_prims", _lib.stream_handle(), m.c, _lib.CULL_PT, _lib.ptr(ijt), int(ijt.shape[0]),
                   _lib.ptr(G["pts"].lowers), _lib.ptr(G["pts"].uppers), _lib.ptr(G["tri"].lowers),
                   _lib.ptr(G["tri"].uppers), _lib.ptr(out), cnt.value, ctypes.byref(cnt))
    assert rc == _lib.ABD_OK
    np.testing.assert_array_equal(N(out)[:len(full)], full)


def test_trajectory_boxes_bit_exact(lib):
    from abd_hip import _lib
    from abd_hip.culling import BvhBuilder
    spec = contact_scene()
    m = _gpu_scene(spec)
    sc = O.OracleScene(spec)
    rng = np.random.default_rng(5)
    x0 = O.project(sc, spec.A, spec.p)
    x1 = x0 + 0.01 * rng.normal(size=x0.shape)
    m.x.copy_(T(x0))
    m.xk.copy_(T(x1))
    bb = BvhBuilder(m)
    for kind, fn, dil, off in ((O.KIND_POINT, "build_point_bvh", 1e-3, None), (O.KIND_EDGE, "build_edge_bvh", 5e-4, m.edges),
                               (O.KIND_TRI, "bulid_triangle_bvh", 0.0, m.triangles)):
        args = (m.x, dil) if off is None else (m.x, off, dil)
        b0 = getattr(bb, fn)(*args)
        b1 = getattr(bb, fn)(*args)
        if kind == O.KIND_POINT:
            bb.point_bvh_to_traj(m.xk, b0, b1)
        elif kind == O.KIND_EDGE:
            bb.edge_bvh_to_traj(m.xk, m.edges, b0, b1)
        else:
            bb.triangle_bvh_to_traj(m.xk, m.triangles, b0, b1)
        lo, hi = O.prim_boxes(sc, kind, x0, dil)
        lo, hi = O.prim_boxes_traj(sc, kind, x1, lo, hi)
        np.testing.assert_array_equal(N(b1.lowers), lo)
        np.testing.assert_array_equal(N(b1.uppers), hi)
    bod = bb.build_body_bvh(m, 5e-4)
    trj = bb.build_body_bvh(m, 5e-4)
    bb.body_bvh_to_traj(m, bod, trj)
    lo, hi = O.body_boxes(sc, x0, 5e-4)
    lo, hi = O.body_boxes_traj(sc, x1, lo, hi)
    np.testing.assert_array_equal(N(trj.lowers), lo)
    np.testing.assert_array_equal(N(trj.uppers), hi)


# --------------------------------------------------------------- energy and gh
@pytest.mark.parametrize("psd", [0, 1, 2, 3])
def test_energy_and_gh_match_oracle(lib, psd):
    from abd_hip.ipc import IPCContactEnergy
    spec = contact_scene()
    cp = params(ground=spec.ground, psd=psd)
    m = _gpu_scene(spec)
    sc = O.OracleScene(spec)
    x = O.project(sc, spec.A, spec.p)
    m.x.copy_(T(x))
    m.xk.copy_(T(x))
    G = _stage_lists(m, cp)
    R = oracle_lists(sc, x, cp.dhat, cp.ground)
    prm = O.params(cp)
    ipc = IPCContactEnergy(cp)
    E = ipc.energy([G["ee"], G["pt"], G["vg"], m])
    E_ref = O.contact_energy(sc, prm, x, R["pt"], R["ee"], R["vg"])
    assert E_ref > 0
    assert abs(E - E_ref) <= 1e-9 * abs(E_ref)
    nb, nij = spec.n_bodies, len(R["ij"])
    rng = np.random.default_rng(1)
    g0 = rng.normal(size=(4 * nb, 3))
    b0 = rng.normal(size=((nb + 2 * nij) * 16, 3, 3))
    g, blocks = T(g0), T(b0)
    hl, nact = ipc.gh([G["pt"], G["ee"], G["vg"], nij, m, g, blocks], return_active=True)
    g_ref, b_ref, hl_ref = O.contact_gh(sc, prm, x, R["pt"], R["ee"], R["vg"], nij, g0, b0)
    assert nact > 0
    np.testing.assert_array_equal(N(ipc.highlight_bits), hl_ref)
    np.testing.assert_array_equal(N(hl), (hl_ref & 2) != 0)
    assert rel_err(N(g) - g0, g_ref - g0) <= 1e-9
    assert rel_err(N(blocks) - b0, b_ref - b0) <= 1e-9


def test_gh_rejects_unsorted_lists(lib):
    from abd_hip._lib import AbdError
    from abd_hip.ipc import IPCContactEnergy
    spec = contact_scene()
    cp = params(ground=spec.ground)
    m = _gpu_scene(spec)
    st = m.states()
    from abd_hip import _lib
    _lib.call("abd_project_vertices", _lib.stream_handle(), m.c, _lib.ptr(st.A), _lib.ptr(st.p), _lib.ptr(m.xk))
    m.x.copy_(m.xk)
    G = _stage_lists(m, cp)
    ee = G["ee"]
    assert G["pairs"].nij >= 2 and ee.shape[0] >= 2
    bad = torch.flip(ee, dims=[0]).contiguous()
    if bool((bad[1:, 2] >= bad[:-1, 2]).all()):
        pytest.skip("all EE candidates belong to one body pair")
    nb, nij = spec.n_bodies, G["pairs"].nij
    g = torch.zeros((4 * nb, 3), dtype=torch.float64, device=DEV)
    blocks = torch.zeros(((nb + 2 * nij) * 16, 3, 3), dtype=torch.float64, device=DEV)
    with pytest.raises(AbdError) as e:
        IPCContactEnergy(cp).gh([G["pt"], bad, G["vg"], nij, m, g, blocks])
    assert e.value.code == _lib.ABD_ERR_UNSORTED


@pytest.mark.parametrize("psd", [-1, 4])
def test_entry_points_reject_unknown_psd(lib, psd):
    """psd outside 0..3 is an argument error of every entry point that reads it."""
    from abd_hip import _lib
    from abd_hip._lib import AbdError
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.ipc import IPCContactEnergy, pair_gh
    spec = contact_scene()
    good = params(ground=spec.ground)
    bad = good.to_c()
    bad.psd = psd  # past ContactParams' own check: the library's
    cp = params(ground=spec.ground)
    cp.to_c = lambda: bad
    x4 = torch.zeros((1, 4, 3), dtype=torch.float64, device=DEV)
    with pytest.raises(AbdError) as e:
        pair_gh(0, cp, x4, x4.clone())
    assert e.value.code == _lib.ABD_ERR_ARG
    m = _gpu_scene(spec)
    _lib.call("abd_project_vertices", _lib.stream_handle(), m.c, _lib.ptr(m.states().A), _lib.ptr(m.states().p),
              _lib.ptr(m.xk))
    m.x.copy_(m.xk)
    G = _stage_lists(m, good)
    nb, nij = spec.n_bodies, G["pairs"].nij
    g = torch.zeros((4 * nb, 3), dtype=torch.float64, device=DEV)
    blocks = torch.zeros(((nb + 2 * nij) * 16, 3, 3), dtype=torch.float64, device=DEV)
    with pytest.raises(AbdError) as e:
        IPCContactEnergy(cp).gh([G["pt"], G["ee"], G["vg"], nij, m, g, blocks])
    assert e.value.code == _lib.ABD_ERR_ARG
    with pytest.raises(AbdError) as e:
        AffineBodySimulator(spec, cp, "cuda", native=True)
    assert e.value.code == _lib.ABD_ERR_ARG


def test_inertia_matches_oracle(lib):
    from abd_hip.orthogonal_energy import InertialEnergy
    from abd_hip import scenes
    spec = scenes.config_medium(seed=1, n_side=2)
    cp = params(ground=spec.ground)
    m = _gpu_scene(spec)
    st = m.states()
    rng = np.random.default_rng(2)
    A = spec.A + 0.01 * rng.normal(size=spec.A.shape)
    p = spec.p + 0.01 * rng.normal(size=spec.p.shape)
    Adot = 0.1 * rng.normal(size=spec.A.shape)
    for name, v in (("A", A), ("Ak", A), ("p", p), ("pk", p), ("Adot", Adot)):
        getattr(st, name).copy_(T(v))
    prm = O.params(cp)
    ie = InertialEnergy(cp, m)
    E = ie.energy(st)
    E_ref = O.inertia_energy(prm, A, p, spec.A, spec.p, Adot, spec.pdot, spec.mass, spec.I0)
    assert abs(E - E_ref) <= 1e-9 * abs(E_ref)
    g = torch.full((4 * spec.n_bodies, 3), 7.0, dtype=torch.float64, device=DEV)
    ie.gradient(g, st)
    g_ref = O.inertia_gradient(prm, A, p, spec.A, spec.p, Adot, spec.pdot, spec.mass, spec.I0)
    assert rel_err(N(g), g_ref) <= 1e-12
    b0 = rng.normal(size=(16 * spec.n_bodies, 3, 3))
    blocks = T(b0)
    ie.hessian(blocks, st)
    b_ref = O.inertia_hessian(prm, A, b0, spec.mass, spec.I0)
    assert rel_err(N(blocks), b_ref) <= 1e-12


# ---------------------------------------------------------------------- solver
def _random_system(nb, ij, rng, spd=True):
    nij = len(ij)
    nslot = nb + 2 * nij
    blocks = np.zeros((nslot * 16, 3, 3))
    dense = np.zeros((12 * nb, 12 * nb))
    def put(slot, M):
        for ii in range(4):
            for jj in range(4):
                blocks[16 * slot + ii + 4 * jj] = M[3 * ii:3 * ii + 3, 3 * jj:3 * jj + 3]
    offs = []
    for k, (i, j) in enumerate(ij):
        C = 0.3 * rng.normal(size=(12, 12))
        put(nb + k, C)
        put(nb + nij + k, C.T)
        dense[12 * i:12 * i + 12, 12 * j:12 * j + 12] += C
        dense[12 * j:12 * j + 12, 12 * i:12 * i + 12] += C.T
        offs.append(C)
    for b in range(nb):
        Q = rng.normal(size=(12, 12))
        D = Q @ Q.T + 12.0 * np.eye(12) * (1 + len(ij))
        if not spd:
            D = D + 0.5 * rng.normal(size=(12, 12))
        put(b, D)
        dense[12 * b:12 * b + 12, 12 * b:12 * b + 12] += D
    return blocks, dense


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("nb", [40, 300])
def test_bsr_solve_matches_oracle(lib, method, nb):
    """method 2 is CG with the 12x12 block-Jacobi preconditioner (beyond the reference)."""
    """nb = 300 exercises the large-system paths (CG over several workgroups; BiCGSTAB
    with its vectors outside LDS)."""
    from abd_hip import _lib
    from abd_hip.sparse import BsrMatrix, SolveInfo, bicgstab, bsr_cg
    rng = np.random.default_rng(3 + method + nb)
    pairs = sorted({tuple(sorted(rng.choice(nb, 2, replace=False))) for _ in range(7 * nb // 4)})
    ij = np.array(pairs, np.int32)
    blocks, dense = _random_system(nb, ij, rng, spd=True)
    b = rng.normal(size=(4 * nb, 3))
    x0 = 0.01 * rng.normal(size=(4 * nb, 3))
    ji = ij[np.lexsort((ij[:, 0], ij[:, 1]))]
    nij = len(ij)
    row_ptr = torch.empty(nb + 1, dtype=torch.int32, device=DEV)
    cols = torch.empty(nb + 2 * nij, dtype=torch.int32, device=DEV)
    slots = torch.empty(nb + 2 * nij, dtype=torch.int32, device=DEV)
    tij, tji = T(ij), T(ji)
    _lib.call("abd_bsr_structure", _lib.stream_handle(), nb, _lib.ptr(tij), _lib.ptr(tji), nij,
              _lib.ptr(row_ptr), _lib.ptr(cols), _lib.ptr(slots))
    # the CSR: columns ascending, every slot once
    rp, cl, sl = N(row_ptr), N(cols), N(slots)
    assert rp[-1] == nb + 2 * nij and sorted(sl.tolist()) == list(range(nb + 2 * nij))
    for r in range(nb):
        assert np.all(np.diff(cl[rp[r]:rp[r + 1]]) > 0)
    A = BsrMatrix(nb, row_ptr, cols, slots, T(blocks))
    x = T(x0)
    info = SolveInfo(DEV)
    tol = 1e-10
    if method == 0:
        bicgstab(A, T(b), x, tol, 0, 10, info)
        x_ref, it_ref, _ = O.bicgstab(nb, ij, blocks, b, x0, tol)
    elif method == 1:
        bsr_cg(A, x, T(b), 0, tol, 10, info)
        x_ref, it_ref, _ = O.bsr_cg(nb, ij, blocks, b, x0, tol)
    else:
        bsr_cg(A, x, T(b), 0, tol, 10, info, block_precond=True)
        x_ref, it_ref, _ = O.bsr_cg_block(nb, ij, blocks, b, x0, tol)
    xs = np.linalg.solve(dense, b.reshape(-1)).reshape(-1, 3)
    assert rel_err(N(x), xs) <= 1e-7
    assert rel_err(N(x), x_ref) <= 1e-7
    assert abs(int(N(info.iters)[0]) - it_ref) <= 10


# -------------------------------------------------------------------------- CCD
@pytest.mark.parametrize("nb,npair", [(5, 0), (12, 20), (300, 900)])
def test_set_triplets_equal_oracle(lib, nb, npair):
    """abd_set_triplets (orthogonal_energy.py:138-165 over all Nb + 2 Nij slots) equals the
    oracle's triplets, which assemble the dense matrix (test_oracle_stages)."""
    from abd_hip.orthogonal_energy import set_triplets
    rng = np.random.default_rng(nb + npair)
    pairs = sorted({tuple(sorted(rng.choice(nb, 2, replace=False))) for _ in range(npair)})
    ij = np.array(pairs, np.int32).reshape(-1, 2)
    nij = ij.shape[0]
    rows = torch.full(((nb + 2 * nij) * 16,), -7, dtype=torch.int32, device=DEV)
    cols = torch.full_like(rows, -7)
    set_triplets(nb, nij, torch.from_numpy(ij).to(DEV), rows, cols)
    r_ref, c_ref = O.set_triplets(nb, ij)
    np.testing.assert_array_equal(rows.cpu().numpy(), r_ref)
    np.testing.assert_array_equal(cols.cpu().numpy(), c_ref)


def test_cubic_roots_match_oracle(lib):
    from abd_hip.ccd import cubic_roots
    rng = np.random.default_rng(7)
    coefs = []
    for _ in range(400):
        r = np.sort(rng.uniform(-0.2, 1.2, 3))
        a0 = rng.uniform(0.1, 5.0) * (1 if rng.uniform() < 0.5 else -1)
        coefs.append(np.array([-r[0] * r[1] * r[2], r[0] * r[1] + r[0] * r[2] + r[1] * r[2], -r.sum(), 1.0]) * a0)
    coefs += [np.array([0.3, -1.0, 0.0, 0.0]), np.array([0.1, 0.0, -1.0, 0.0]), np.zeros(4),
              np.array([1.0, 2.0, 3.0, 4.0])]
    C = np.array(coefs)
    nr, roots = cubic_roots(T(C), 0.0, 1.0)
    nr, roots = N(nr), N(roots)
    for i, c in enumerate(C):
        n_ref, r_ref = O.cubic_roots(c, 0.0, 1.0)
        assert nr[i] == n_ref, i
        np.testing.assert_allclose(roots[i][:n_ref], r_ref[:n_ref], rtol=0, atol=1e-12)


@pytest.mark.parametrize("kind", [0, 1])
def test_collision_time_matches_oracle(lib, kind):
    from abd_hip.ccd import collision_time
    rng = np.random.default_rng(8 + kind)
    P0, P1 = [], []
    for i in range(300):
        x4 = near_contact_pt(rng, 0.05) if kind == 0 else near_contact_ee(rng, 0.05)
        v = 0.1 * rng.normal(size=(4, 3))
        if i % 3 == 0:
            v[0] -= 0.08 * np.cross(x4[2] - x4[1], x4[3] - x4[1])  # drive through
        P0.append(x4)
        P1.append(x4 + v)
    P0, P1 = np.array(P0), np.array(P1)
    t = N(collision_time(kind, T(P0), T(P1)))
    hits = 0
    for i in range(len(P0)):
        t_ref = O.collision_time(kind, P0[i], P1[i])
        assert abs(t[i] - t_ref) <= 1e-12, (i, t[i], t_ref)
        hits += t_ref < 1.0
    assert hits > 10


def test_scene_toi_matches_oracle(lib):
    from abd_hip.ccd import toi
    from abd_hip.culling import BvhBuilder, cull, cull_vg, ij_list
    spec = contact_scene()
    cp = params(ground=spec.ground)
    m = _gpu_scene(spec)
    sc = O.OracleScene(spec)
    x0 = O.project(sc, spec.A, spec.p)
    rng = np.random.default_rng(11)
    # move the bodies rigidly toward each other / the ground
    p1 = spec.p + np.array([[0, -0.002, 0], [0, -0.004, 0], [0, -0.003, 0], [-0.003, -0.001, 0]])
    x1 = O.project(sc, spec.A, p1)
    m.x.copy_(T(x0))
    m.xk.copy_(T(x1))
    bb = BvhBuilder(m)
    dh = 0.5 * cp.dhat
    b = bb.build_body_bvh(m, dh)
    bt = bb.build_body_bvh(m, dh)
    bb.body_bvh_to_traj(m, b, bt)
    tri, edg, pts = (bb.bulid_triangle_bvh(m.x, m.triangles, 0.0), bb.build_edge_bvh(m.x, m.edges, dh),
                     bb.build_point_bvh(m.x, cp.dhat))
    trit, edgt, ptst = (bb.bulid_triangle_bvh(m.x, m.triangles, 0.0), bb.build_edge_bvh(m.x, m.edges, dh),
                        bb.build_point_bvh(m.x, cp.dhat))
    bb.triangle_bvh_to_traj(m.xk, m.triangles, tri, trit)
    bb.edge_bvh_to_traj(m.xk, m.edges, edg, edgt)
    bb.point_bvh_to_traj(m.xk, pts, ptst)
    pairs = ij_list(bt, m)
    ee = cull(pairs, edgt)
    pt = cull(pairs, ptst, trit)
    vg = cull_vg(bt.lowers, ptst, m, cp.dhat, cp.ground)
    t = toi(m, cp, pt, ee, vg)
    t_ref = O.toi(sc, O.params(cp), x0, x1, N(pt), N(ee), N(vg))
    assert t_ref < 1.0
    assert abs(t - t_ref) <= 1e-12


def test_dot_and_updates(lib):
    from abd_hip import _lib
    rng = np.random.default_rng(12)
    a, b = rng.normal(size=10001), rng.normal(size=10001)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    ta, tb = T(a), T(b)  # keep both alive: a freed temporary's block is handed to the next one
    _lib.call("abd_dot", _lib.stream_handle(), _lib.ptr(ta), _lib.ptr(tb), a.size, _lib.ptr(out))
    assert abs(N(out)[0] - np.dot(a, b)) <= 1e-12 * np.abs(a * b).sum()
    nb = 5
    A, p, dq = rng.normal(size=(nb, 3, 3)), rng.normal(size=(nb, 3)), rng.normal(size=(4 * nb, 3))
    Ao, po = torch.zeros((nb, 3, 3), dtype=torch.float64, device=DEV), torch.zeros((nb, 3), dtype=torch.float64,
                                                                                  device=DEV)
    tA, tp, tdq = T(A), T(p), T(dq)
    _lib.call("abd_update_q", _lib.stream_handle(), nb, _lib.ptr(tA), _lib.ptr(tp), _lib.ptr(tdq), 0.37,
              _lib.ptr(Ao), _lib.ptr(po))
    A_ref, p_ref = np.zeros_like(A), np.zeros_like(p)
    O.lib().orc_update_q(nb, A.ctypes.data, p.ctypes.data, dq.ctypes.data, 0.37, A_ref.ctypes.data, p_ref.ctypes.data)
    np.testing.assert_array_equal(N(Ao), A_ref)
    np.testing.assert_array_equal(N(po), p_ref)


# ---------------------------------------------------------------- fixtures
def test_gpu_matches_golden_fixtures(lib):
    import os
    from abd_hip.ipc import pair_gh
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixed_pair.npz"))
    cp = params(dhat=float(f["dhat"]))
    for kind, tag in ((0, "pt"), (1, "ee")):
        X = np.concatenate([f["x"][None], f[f"{tag}_act_x"]])
        R = np.concatenate([f["x"][None], f[f"{tag}_act_r"]])
        act, g24, H24 = pair_gh(kind, cp, T(X), T(R))
        act, g24, H24 = N(act), N(g24), N(H24)
        assert not act[0] and act[1:].all()
        for i in range(1, len(X)):
            assert rel_err(g24[i], f[f"{tag}_act_g24"][i - 1]) <= 1e-9
            assert rel_err(H24[i], f[f"{tag}_act_H24"][i - 1]) <= 1e-9


@pytest.mark.parametrize("psd", [1, 2])
@pytest.mark.parametrize("kind", [0, 1])
def test_pair_gh_psd_mode_matches_oracle(lib, kind, psd):
    """psd=1 (contact-space PSD projection) and psd=2 (exact Hessian) against the oracle."""
    from abd_hip.ipc import pair_gh
    rng = np.random.default_rng(20 + kind)
    cp = params(psd=psd)
    X, R = [], []
    for i in range(128):
        d = rng.uniform(1e-5, 9.9e-4) * (1 if i % 2 else -1)
        x4 = near_contact_pt(rng, d) if kind == 0 else near_contact_ee(rng, d)
        X.append(x4)
        R.append(rest_of(x4, rng))
    X, R = np.array(X), np.array(R)
    act, g24, H24 = pair_gh(kind, cp, T(X), T(R))
    act, g24, H24 = N(act), N(g24), N(H24)
    prm = O.params(cp)
    for i in range(len(X)):
        a, s, g, Hi, Hj, Hij = O.pair_gh(kind, prm, X[i], R[i])
        assert bool(act[i]) == a
        if a:
            Href = np.block([[Hi, Hij], [Hij.T, Hj]])
            assert rel_err(H24[i], Href) <= 1e-9
            assert rel_err(g24[i], g) <= 1e-9


@pytest.mark.parametrize("which", [0, 1, 2])
def test_cull_keep_equals_culling(lib, which):
    """abd_cull_keep on the lists of larger boxes (5 dhat) reproduces the culling of the
    current boxes exactly: pairs in both orders and every tuple, in order."""
    import ctypes as C
    from abd_hip import _lib
    spec = _scenes()[which]
    cp = params(ground=spec.ground)
    m = _gpu_scene(spec)
    st = m.states()
    _lib.call("abd_project_vertices", _lib.stream_handle(), m.c, _lib.ptr(st.A), _lib.ptr(st.p), _lib.ptr(m.xk))
    G = _stage_lists(m, cp)
    W = _stage_lists(m, params(ground=spec.ground, dhat=5 * cp.dhat))
    nij, npt, nee, nvg = len(W["pairs"].ij), len(W["pt"]), len(W["ee"]), len(W["vg"])
    assert nij >= len(G["pairs"].ij) and npt + nee >= len(G["pt"]) + len(G["ee"])
    out = dict(ij=torch.zeros((nij + 1, 2), dtype=torch.int32, device=DEV),
               ji=torch.zeros((nij + 1, 2), dtype=torch.int32, device=DEV),
               pt=torch.zeros((npt + 1, 5), dtype=torch.int32, device=DEV),
               ee=torch.zeros((nee + 1, 5), dtype=torch.int32, device=DEV),
               vg=torch.zeros((nvg + 1, 2), dtype=torch.int32, device=DEV))
    counts = (C.c_int64 * 5)()
    P = _lib.ptr
    _lib.call("abd_cull_keep", _lib.stream_handle(), m.c, P(W["pairs"].ij), P(W["pairs"].ji), nij, P(W["pt"]), npt,
              P(W["ee"]), nee, P(W["vg"]), nvg, P(G["bodies"].lowers), P(G["bodies"].uppers), P(G["pts"].lowers),
              P(G["pts"].uppers), P(G["edg"].lowers), P(G["edg"].uppers), P(G["tri"].lowers), P(G["tri"].uppers),
              C.c_double(cp.dhat), C.c_double(cp.ground), P(out["ij"]), P(out["ji"]), P(out["pt"]), P(out["ee"]),
              P(out["vg"]), counts)
    n = list(counts)
    np.testing.assert_array_equal(N(out["ij"])[:n[0]], N(G["pairs"].ij))
    np.testing.assert_array_equal(N(out["ji"])[:n[1]], N(G["pairs"].ji))
    np.testing.assert_array_equal(N(out["pt"])[:n[2]], N(G["pt"]).reshape(-1, 5))
    np.testing.assert_array_equal(N(out["ee"])[:n[3]], N(G["ee"]).reshape(-1, 5))
    np.testing.assert_array_equal(N(out["vg"])[:n[4]], N(G["vg"]).reshape(-1, 2))
