"""Candidate culling (culling.py / culling2.py in the reference).

`BvhBuilder` keeps the reference's method names (culling2.py:83-210) but
builds box sets over the whole flattened scene instead of one Warp BVH per body;
each returned `BoxSet` exposes `.lowers` / `.uppers` (float32 [n,3]) like a
`wp.Bvh`, and `.refit()` is a no-op because culling tests the boxes directly
(abd_cull_prims: exact hull filter + exhaustive chunked tests).

`cull(ij_list, bvh_list1, bvh_list2=None)` (culling.py:8) and
`cull_vg(lowers, bvh_list, meshes)` (culling.py:54) return int32 tuple lists
[N,5] / [N,2] like the reference's compressed vec5i / vec2i arrays.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional

import torch

from . import _lib
from .affine_body import Meshes


@dataclasses.dataclass
class BoxSet:
    """FP32 boxes of one primitive kind for every body of a scene."""
    meshes: Meshes
    kind: int
    lowers: torch.Tensor
    uppers: torch.Tensor

    def refit(self):  # the wp.Bvh API; boxes are tested directly
        return self


@dataclasses.dataclass
class BodyPairs:
    """Output of ij_list(): pairs sorted by (i, j), the same pairs by (j, i), and
    the block-row structure of the Newton matrix built from them."""
    ij: torch.Tensor
    ji: torch.Tensor
    row_ptr: torch.Tensor
    cols: torch.Tensor
    slots: torch.Tensor

    @property
    def nij(self) -> int:
        return int(self.ij.shape[0])

    @property
    def shape(self):
        return self.ij.shape

    def numpy(self):
        return self.ij.cpu().numpy()


def _stream():
    return _lib.stream_handle()


class BvhBuilder:
    """culling2.py:83-210.  The positions V passed in are scene-wide [V,3] tensors
    (meshes.x / meshes.xk), the topology the scene's flattened triangles / edges."""

    def __init__(self, meshes: Meshes):
        self.meshes = meshes

    def _new(self, kind: int, n: int) -> BoxSet:
        dev = self.meshes.device
        return BoxSet(self.meshes, kind, torch.zeros((n, 3), dtype=torch.float32, device=dev),
                      torch.zeros((n, 3), dtype=torch.float32, device=dev))

    # points (culling2.py:90-109)
    def build_point_bvh(self, V, dialation) -> BoxSet:
        bvh = self._new(_lib.KIND_POINT, self.meshes.n_verts)
        return self.update_point_bvh(V, dialation, bvh)

    def update_point_bvh(self, V, dialation, bvh: BoxSet) -> BoxSet:
        _lib.call("abd_prim_boxes", _stream(), self.meshes.c, _lib.KIND_POINT, _lib.ptr(V), float(dialation),
                  _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def point_bvh_to_traj(self, V, bvh_point: BoxSet, bvh_traj: BoxSet) -> BoxSet:
        bvh_traj.lowers.copy_(bvh_point.lowers)
        bvh_traj.uppers.copy_(bvh_point.uppers)
        _lib.call("abd_prim_boxes_traj", _stream(), self.meshes.c, _lib.KIND_POINT, _lib.ptr(V),
                  _lib.ptr(bvh_traj.lowers), _lib.ptr(bvh_traj.uppers))
        return bvh_traj

    # triangles (culling2.py:111-130; the reference spells it "bulid")
    def bulid_triangle_bvh(self, V, F, dialation) -> BoxSet:
        bvh = self._new(_lib.KIND_TRI, self.meshes.n_tris)
        return self.update_triangle_bvh(V, F, dialation, bvh)

    build_triangle_bvh = bulid_triangle_bvh

    def update_triangle_bvh(self, V, F, dialation, bvh: BoxSet) -> BoxSet:
        _lib.call("abd_prim_boxes", _stream(), self.meshes.c, _lib.KIND_TRI, _lib.ptr(V), float(dialation),
                  _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def triangle_bvh_to_traj(self, V, F, bvh_triangle: BoxSet, bvh_traj: BoxSet) -> BoxSet:
        bvh_traj.lowers.copy_(bvh_triangle.lowers)
        bvh_traj.uppers.copy_(bvh_triangle.uppers)
        _lib.call("abd_prim_boxes_traj", _stream(), self.meshes.c, _lib.KIND_TRI, _lib.ptr(V),
                  _lib.ptr(bvh_traj.lowers), _lib.ptr(bvh_traj.uppers))
        return bvh_traj

    # edges (culling2.py:132-152)
    def build_edge_bvh(self, V, E, dialation) -> BoxSet:
        bvh = self._new(_lib.KIND_EDGE, self.meshes.n_edges)
        return self.update_edge_bvh(V, E, dialation, bvh)

    def update_edge_bvh(self, V, E, dialation, bvh: BoxSet) -> BoxSet:
        _lib.call("abd_prim_boxes", _stream(), self.meshes.c, _lib.KIND_EDGE, _lib.ptr(V), float(dialation),
                  _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def edge_bvh_to_traj(self, V, E, bvh_edge: BoxSet, bvh_traj: BoxSet) -> BoxSet:
        bvh_traj.lowers.copy_(bvh_edge.lowers)
        bvh_traj.uppers.copy_(bvh_edge.uppers)
        _lib.call("abd_prim_boxes_traj", _stream(), self.meshes.c, _lib.KIND_EDGE, _lib.ptr(V),
                  _lib.ptr(bvh_traj.lowers), _lib.ptr(bvh_traj.uppers))
        return bvh_traj

    # bodies (culling2.py:154-210): reduced on the GPU instead of per-body numpy
    def build_body_bvh(self, bodies: Meshes, dialation, field: str = "x") -> BoxSet:
        bvh = self._new(-1, bodies.n_bodies)
        _lib.call("abd_body_boxes", _stream(), bodies.c, _lib.ptr(getattr(bodies, field)), float(dialation),
                  _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def update_body_bvh(self, bodies: Meshes, dialation, bvh: BoxSet) -> BoxSet:
        _lib.call("abd_body_boxes", _stream(), bodies.c, _lib.ptr(bodies.xk), float(dialation),
                  _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def body_bvh_to_traj(self, bodies: Meshes, bvh_body: BoxSet, bvh_traj: BoxSet) -> BoxSet:
        bvh_traj.lowers.copy_(bvh_body.lowers)
        bvh_traj.uppers.copy_(bvh_body.uppers)
        _lib.call("abd_body_boxes_traj", _stream(), bodies.c, _lib.ptr(bodies.xk), _lib.ptr(bvh_traj.lowers),
                  _lib.ptr(bvh_traj.uppers))
        return bvh_traj


def ij_list(bvh_bodies: BoxSet, meshes: Optional[Meshes] = None) -> BodyPairs:
    """abd.py:163-168 + culling2.py:214-227: overlapping body pairs (i < j), plus the
    sparse structure of the Newton matrix over them."""
    m = meshes or bvh_bodies.meshes
    dev = m.device
    cap = max(16, getattr(m, "_ij_cap", 16))
    s = _stream()
    while True:
        ij = torch.empty((cap, 2), dtype=torch.int32, device=dev)
        ji = torch.empty((cap, 2), dtype=torch.int32, device=dev)
        cnt = ctypes.c_int64(0)
        rc = _lib.call("abd_cull_bodies", s, m.c, _lib.ptr(bvh_bodies.lowers), _lib.ptr(bvh_bodies.uppers),
                       _lib.ptr(ij), _lib.ptr(ji), cap, ctypes.byref(cnt), ok=(_lib.ABD_OK, _lib.ABD_ERR_OVERFLOW))
        n = cnt.value
        if rc == _lib.ABD_OK:
            break
        cap = 2 * n
    m._ij_cap = max(cap, 2 * n)
    ij, ji = ij[:n], ji[:n]
    nb = m.n_bodies
    row_ptr = torch.empty(nb + 1, dtype=torch.int32, device=dev)
    cols = torch.empty(nb + 2 * n, dtype=torch.int32, device=dev)
    slots = torch.empty(nb + 2 * n, dtype=torch.int32, device=dev)
    _lib.call("abd_bsr_structure", s, nb, _lib.ptr(ij), _lib.ptr(ji), n, _lib.ptr(row_ptr), _lib.ptr(cols),
              _lib.ptr(slots))
    pairs = BodyPairs(ij, ji, row_ptr, cols, slots)
    m.pairs = pairs
    return pairs


def _cull_prims(mode: int, ij: BodyPairs, s1: BoxSet, s2: BoxSet) -> torch.Tensor:
    m = s1.meshes
    key = "_cap_pt" if mode == _lib.CULL_PT else "_cap_ee"
    cap = max(64, getattr(m, key, 64))
    s = _stream()
    ijt = ij.ij if isinstance(ij, BodyPairs) else ij
    nij = int(ijt.shape[0])
    while True:
        out = torch.empty((cap, 5), dtype=torch.int32, device=m.device)
        cnt = ctypes.c_int64(0)
        rc = _lib.call("abd_cull_prims", s, m.c, mode, _lib.ptr(ijt), nij, _lib.ptr(s1.lowers), _lib.ptr(s1.uppers),
                       _lib.ptr(s2.lowers), _lib.ptr(s2.uppers), _lib.ptr(out), cap, ctypes.byref(cnt),
                       ok=(_lib.ABD_OK, _lib.ABD_ERR_OVERFLOW))
        n = cnt.value
        if rc == _lib.ABD_OK:
            break
        cap = n + n // 2 + 64
    setattr(m, key, max(cap, n + n // 4 + 64))
    return out[:n]


def cull(ij_list, _bvh_list1: BoxSet, _bvh_list2: Optional[BoxSet] = None) -> torch.Tensor:
    """culling.py:8 / :32-52 (cull_ semantics, defect D1 fixed).  With a second box
    set: PT candidates (points of set 1 vs triangles of set 2, both body orders);
    without: EE candidates."""
    if _bvh_list2 is None:
        return _cull_prims(_lib.CULL_EE, ij_list, _bvh_list1, _bvh_list1)
    return _cull_prims(_lib.CULL_PT, ij_list, _bvh_list1, _bvh_list2)


def cull_vg(lowers: torch.Tensor, _bvh_list: BoxSet, bodies: Meshes, dhat: float, ground: float) -> torch.Tensor:
    """culling.py:54-76: vertex-ground candidates (body, vertex)."""
    m = bodies
    cap = max(64, getattr(m, "_cap_vg", 64))
    s = _stream()
    while True:
        out = torch.empty((cap, 2), dtype=torch.int32, device=m.device)
        cnt = ctypes.c_int64(0)
        rc = _lib.call("abd_cull_vg", s, m.c, _lib.ptr(lowers), _lib.ptr(_bvh_list.lowers), _lib.ptr(_bvh_list.uppers),
                       float(dhat), float(ground), _lib.ptr(out), cap, ctypes.byref(cnt),
                       ok=(_lib.ABD_OK, _lib.ABD_ERR_OVERFLOW))
        n = cnt.value
        if rc == _lib.ABD_OK:
            break
        cap = n + n // 2 + 64
    m._cap_vg = max(cap, n + n // 4 + 64)
    return out[:n]
