"""abd_toi_noalloc (the allocation-free form of abd_toi) against abd_toi through
the C ABI: the same TOI bit for bit on a swept candidate set of the 256-scene
batch (every scene's bodies moved along their initial velocities), on two calls in a row (the second with shorter lists, so the fixed
scratch holds the first call's chunk partials when the second starts), and the
capacity error for lists beyond the fixed scratch (returned before any launch)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _swept_set(spec, cp, m, x1):
    from abd_hip.culling import BvhBuilder, cull, cull_vg, ij_list
    m.xk.copy_(x1)
    bb = BvhBuilder(m)
    dh = 0.5 * cp.dhat
    b = bb.build_body_bvh(m, dh)
    bt = bb.build_body_bvh(m, dh)
    bb.body_bvh_to_traj(m, b, bt)
    trit, edgt, ptst = (bb.bulid_triangle_bvh(m.x, m.triangles, 0.0), bb.build_edge_bvh(m.x, m.edges, dh),
                        bb.build_point_bvh(m.x, cp.dhat))
    tri, edg, pts = (bb.bulid_triangle_bvh(m.x, m.triangles, 0.0), bb.build_edge_bvh(m.x, m.edges, dh),
                     bb.build_point_bvh(m.x, cp.dhat))
    bb.triangle_bvh_to_traj(m.xk, m.triangles, tri, trit)
    bb.edge_bvh_to_traj(m.xk, m.edges, edg, edgt)
    bb.point_bvh_to_traj(m.xk, pts, ptst)
    pairs = ij_list(bt, m)
    return cull(pairs, ptst, trit), cull(pairs, edgt), cull_vg(bt.lowers, ptst, m, cp.dhat, cp.ground)


def _toi(name, m, pc, pt, ee, vg):
    from abd_hip import _lib
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    rc = getattr(_lib.load(), name)(_lib.stream_handle(), ctypes.byref(m.c), ctypes.byref(pc), _lib.ptr(m.x),
                                    _lib.ptr(m.xk), _lib.ptr(pt), int(pt.shape[0]), _lib.ptr(ee), int(ee.shape[0]),
                                    _lib.ptr(vg), int(vg.shape[0]), _lib.ptr(out))
    assert rc == 0, (name, rc)
    torch.cuda.synchronize()
    return float(out.item())


def test_noalloc_toi_matches_abd_toi():
    import abd_hip
    from abd_hip import _lib, scenes
    from abd_hip.affine_body import Meshes
    from abd_hip.const_params import ContactParams
    abd_hip.load()
    spec = scenes.config_batched()
    cp = ContactParams(ground=spec.ground)
    pc = cp.to_c()
    m = Meshes(spec, "cuda")
    st = m.states()
    _lib.call("abd_project_vertices", _lib.stream_handle(), ctypes.byref(m.c), _lib.ptr(st.A), _lib.ptr(st.p),
              _lib.ptr(m.x))
    # every body moved by 0.06 of its initial velocity: stacked neighbours and the floor are crossed
    p1 = (st.p + 0.06 * st.pdot).contiguous()
    x1 = torch.empty_like(m.x)
    _lib.call("abd_project_vertices", _lib.stream_handle(), ctypes.byref(m.c), _lib.ptr(st.A), _lib.ptr(p1),
              _lib.ptr(x1))
    pt, ee, vg = _swept_set(spec, cp, m, x1)
    print(f"swept set: {pt.shape[0]} PT, {ee.shape[0]} EE, {vg.shape[0]} VG", flush=True)
    # the first call as large as the fixed scratch takes (the lists' leading parts)
    cap = _lib.TOI_NOALLOC_MAX_CANDIDATES
    k_pt = min(int(pt.shape[0]), cap // 3)
    k_ee = min(int(ee.shape[0]), cap - k_pt - int(vg.shape[0]))
    pt, ee = pt[:k_pt].contiguous(), ee[:k_ee].contiguous()
    n = (int(pt.shape[0]), int(ee.shape[0]), int(vg.shape[0]))
    assert sum(n) <= cap and sum(n) > cap // 2
    t_ref = _toi("abd_toi", m, pc, pt, ee, vg)
    t = _toi("abd_toi_noalloc", m, pc, pt, ee, vg)
    assert t_ref < 1.0
    assert t == t_ref
    # a second, smaller call: the fixed scratch still holds the first call's partials
    k = (n[0] // 5, n[1] // 7, n[2] // 3)
    pt2, ee2, vg2 = pt[:k[0]].contiguous(), ee[:k[1]].contiguous(), vg[:k[2]].contiguous()
    t2_ref = _toi("abd_toi", m, pc, pt2, ee2, vg2)
    t2 = _toi("abd_toi_noalloc", m, pc, pt2, ee2, vg2)
    assert t2 == t2_ref
    # and the first input once more after the smaller one
    assert _toi("abd_toi_noalloc", m, pc, pt, ee, vg) == t_ref


def test_noalloc_toi_rejects_lists_beyond_its_scratch():
    import abd_hip
    from abd_hip import _lib, scenes
    from abd_hip.affine_body import Meshes
    from abd_hip.const_params import ContactParams
    abd_hip.load()
    spec = scenes.config_small()
    cp = ContactParams(ground=spec.ground)
    pc = cp.to_c()
    m = Meshes(spec, "cuda")
    pt = torch.zeros((1, 5), dtype=torch.int32, device="cuda")
    out = torch.ones(1, dtype=torch.float64, device="cuda")
    too_many = _lib.TOI_NOALLOC_MAX_CANDIDATES + 1   # the size check runs before any launch
    rc = _lib.load().abd_toi_noalloc(_lib.stream_handle(), ctypes.byref(m.c), ctypes.byref(pc), _lib.ptr(m.x),
                                     _lib.ptr(m.x), _lib.ptr(pt), too_many, None, 0, None, 0, _lib.ptr(out))
    assert rc == _lib.ABD_ERR_CAPACITY
    assert _lib.error_string(rc).startswith("input exceeds")
    torch.cuda.synchronize()
    assert float(out.item()) == 1.0
