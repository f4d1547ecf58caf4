"""Candidate culling (culling.py / culling2.py in the reference).

`BvhBuilder` keeps the reference's method names (culling2.py:83-210) and takes
either form of their arguments:

- scene-wide: V is a scene array ([V,3] meshes.x / meshes.xk) and the call
  returns / updates a `BoxSet` over every body;
- per body, as abd.py:84-93 and :174-200 call it: V and F are one body's views
  (`meshes[b].xk`, `meshes[b].triangles`) and the call returns / updates that
  body's `BodyBoxes`, a view of one slot of a scene-wide `BoxSet` (the body's
  boxes are computed by the same kernel, so both forms give the same boxes).

Both expose `.lowers` / `.uppers` (float32 [n,3]) like a `wp.Bvh`, and
`.refit()` is a no-op because culling tests the boxes directly (abd_cull_prims:
exact hull filter + exhaustive chunked tests).

`cull(ij_list, bvh_list1, bvh_list2=None)` (culling.py:8) and
`cull_vg(lowers, bvh_list, meshes)` (culling.py:54) take a `BoxSet` or the
per-body list of `BodyBoxes` the builder returned, and return int32 tuple lists
[N,5] / [N,2] like the reference's compressed vec5i / vec2i arrays.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional, Sequence, Union

import numpy as np
import torch

from . import _lib
from .affine_body import Meshes, WarpMesh
from .const_params import ContactParams


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
class BodyBoxes:
    """One body's boxes (the per-body wp.Bvh of culling2.py): views of the body's
    slot of a scene-wide BoxSet."""
    parent: BoxSet
    body: int
    lowers: torch.Tensor
    uppers: torch.Tensor

    def refit(self):
        return self


Boxes = Union[BoxSet, Sequence[BodyBoxes]]


def _scene_boxes(b: Boxes) -> BoxSet:
    """The scene-wide BoxSet behind `b`: b itself, or the parent of a complete per-body
    list (every body once, in body order, all views of one BoxSet)."""
    if isinstance(b, BoxSet):
        return b
    items = list(b)
    if not items or not all(isinstance(x, BodyBoxes) for x in items):
        raise TypeError("expected a BoxSet or the per-body list of BodyBoxes a BvhBuilder returned")
    parent = items[0].parent
    if [x.body for x in items] != list(range(parent.meshes.n_bodies)) or any(x.parent is not parent for x in items):
        raise ValueError("a per-body box list must hold every body of one scene once, in body order")
    return parent


def _scene_meshes(bodies) -> Meshes:
    """The Meshes behind `bodies`: a Meshes, or a sequence of its per-body WarpMesh views."""
    if isinstance(bodies, Meshes):
        return bodies
    items = list(bodies)
    if not items or not all(isinstance(x, WarpMesh) and x.meshes is not None for x in items):
        raise TypeError("expected a Meshes or a sequence of its per-body WarpMesh views")
    m = items[0].meshes
    if [x.id for x in items] != list(range(m.n_bodies)) or any(x.meshes is not m for x in items):
        raise ValueError("a per-body mesh list must hold every body of one scene once, in body order")
    return m


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
    """culling2.py:83-210.  Positions V are a scene array ([V,3] meshes.x / meshes.xk;
    topology the scene's flattened triangles / edges) or one body's view of it (then
    F / E are that body's views and the boxes are BodyBoxes)."""

    def __init__(self, meshes: Optional[Meshes] = None):
        self.meshes = meshes
        # per kind: the scene-wide box sets behind per-body builds, each with the bodies
        # already handed out (the reference builds two lists per kind: current and swept)
        self._parents = {}

    def _scene(self) -> Meshes:
        if self.meshes is None:
            raise ValueError("BvhBuilder needs the scene: BvhBuilder(meshes), or build_body_bvh first")
        return self.meshes

    def _new(self, kind: int, n: int) -> BoxSet:
        dev = self._scene().device
        return BoxSet(self._scene(), kind, torch.zeros((n, 3), dtype=torch.float32, device=dev),
                      torch.zeros((n, 3), dtype=torch.float32, device=dev))

    def _count(self, kind: int) -> int:
        m = self._scene()
        return m.n_verts if kind == _lib.KIND_POINT else m.n_edges if kind == _lib.KIND_EDGE else m.n_tris

    def _range(self, kind: int, b: int):
        sp = self._scene().spec
        off = sp.voff if kind == _lib.KIND_POINT else sp.eoff if kind == _lib.KIND_EDGE else sp.foff
        return int(off[b]), int(off[b + 1])

    def _body_of(self, V: torch.Tensor):
        """(body, scene array) of a per-body position view V, or (None, V) for a scene array."""
        m = self._scene()
        for base in (m.xk, m.x, m.x0, m.x_view):
            if V.data_ptr() == base.data_ptr() and V.shape[0] == m.n_verts:
                return None, base
        for base in (m.xk, m.x, m.x0, m.x_view):
            if V.untyped_storage().data_ptr() != base.untyped_storage().data_ptr():
                continue
            off = (V.data_ptr() - base.data_ptr()) // (3 * V.element_size())
            b = int(np.searchsorted(m.spec.voff, off, side="right")) - 1
            if 0 <= b < m.n_bodies and int(m.spec.voff[b]) == off and int(m.spec.voff[b + 1]) - off == V.shape[0]:
                return b, base
        raise ValueError("V must be a scene position array or one body's view of it")

    def _new_body(self, kind: int, b: int) -> BodyBoxes:
        lst = self._parents.setdefault(kind, [])
        for parent, taken in lst:
            if b not in taken:
                taken.add(b)
                break
        else:
            parent = self._new(kind, self._count(kind))
            lst.append((parent, {b}))
        f0, f1 = self._range(kind, b)
        return BodyBoxes(parent, b, parent.lowers[f0:f1], parent.uppers[f0:f1])

    def _sub_scene(self, kind: int, b: int):
        """The scene struct restricted to body b's primitives of `kind`."""
        m = self._scene()
        c = _lib.AbdScene.from_buffer_copy(m.c)
        f0, f1 = self._range(kind, b)
        if kind == _lib.KIND_POINT:
            c.n_verts = f1 - f0
        elif kind == _lib.KIND_EDGE:
            c.n_edges = f1 - f0
            c.edges = m.edges.data_ptr() + 8 * f0
            c.edge_body = m.edge_body.data_ptr() + 4 * f0
        else:
            c.n_tris = f1 - f0
            c.tris = m.triangles.data_ptr() + 12 * f0
            c.tri_body = m.tri_body.data_ptr() + 4 * f0
        return c

    def _boxes(self, kind: int, V, dialation, bvh):
        if isinstance(bvh, BodyBoxes):
            b, base = self._body_of(V)
            if b != bvh.body:
                raise ValueError(f"positions of body {b} given for the boxes of body {bvh.body}")
            c = self._sub_scene(kind, b)
            x = V if kind == _lib.KIND_POINT else base  # local indices + voff[b] address the scene array
            _lib.call("abd_prim_boxes", _stream(), ctypes.byref(c), kind, _lib.ptr(x), float(dialation),
                      _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        else:
            _lib.call("abd_prim_boxes", _stream(), self._scene().c, kind, _lib.ptr(V), float(dialation),
                      _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def _traj(self, kind: int, V, src, dst):
        dst.lowers.copy_(src.lowers)
        dst.uppers.copy_(src.uppers)
        if isinstance(dst, BodyBoxes):
            b, base = self._body_of(V)
            if b != dst.body:
                raise ValueError(f"positions of body {b} given for the boxes of body {dst.body}")
            c = self._sub_scene(kind, b)
            x = V if kind == _lib.KIND_POINT else base
            _lib.call("abd_prim_boxes_traj", _stream(), ctypes.byref(c), kind, _lib.ptr(x), _lib.ptr(dst.lowers),
                      _lib.ptr(dst.uppers))
        else:
            _lib.call("abd_prim_boxes_traj", _stream(), self._scene().c, kind, _lib.ptr(V), _lib.ptr(dst.lowers),
                      _lib.ptr(dst.uppers))
        return dst

    def _build(self, kind: int, V, dialation):
        b, _ = self._body_of(V)
        bvh = self._new(kind, self._count(kind)) if b is None else self._new_body(kind, b)
        return self._boxes(kind, V, dialation, bvh)

    # points (culling2.py:90-109)
    def build_point_bvh(self, V, dialation):
        return self._build(_lib.KIND_POINT, V, dialation)

    def update_point_bvh(self, V, dialation, bvh):
        return self._boxes(_lib.KIND_POINT, V, dialation, bvh)

    def point_bvh_to_traj(self, V, bvh_point, bvh_traj):
        return self._traj(_lib.KIND_POINT, V, bvh_point, bvh_traj)

    # triangles (culling2.py:111-130; the reference spells it "bulid")
    def bulid_triangle_bvh(self, V, F, dialation):
        return self._build(_lib.KIND_TRI, V, dialation)

    build_triangle_bvh = bulid_triangle_bvh

    def update_triangle_bvh(self, V, F, dialation, bvh):
        return self._boxes(_lib.KIND_TRI, V, dialation, bvh)

    def triangle_bvh_to_traj(self, V, F, bvh_triangle, bvh_traj):
        return self._traj(_lib.KIND_TRI, V, bvh_triangle, bvh_traj)

    # edges (culling2.py:132-152)
    def build_edge_bvh(self, V, E, dialation):
        return self._build(_lib.KIND_EDGE, V, dialation)

    def update_edge_bvh(self, V, E, dialation, bvh):
        return self._boxes(_lib.KIND_EDGE, V, dialation, bvh)

    def edge_bvh_to_traj(self, V, E, bvh_edge, bvh_traj):
        return self._traj(_lib.KIND_EDGE, V, bvh_edge, bvh_traj)

    # bodies (culling2.py:154-210): reduced on the GPU instead of per-body numpy; `bodies`
    # is the Meshes or the sequence of its per-body views (the reference's list of meshes)
    def build_body_bvh(self, bodies, dialation, field: str = "x") -> BoxSet:
        m = _scene_meshes(bodies)
        if self.meshes is None:
            self.meshes = m
        bvh = BoxSet(m, -1, torch.zeros((m.n_bodies, 3), dtype=torch.float32, device=m.device),
                     torch.zeros((m.n_bodies, 3), dtype=torch.float32, device=m.device))
        _lib.call("abd_body_boxes", _stream(), m.c, _lib.ptr(getattr(m, field)), float(dialation),
                  _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def update_body_bvh(self, bodies, dialation, bvh: BoxSet) -> BoxSet:
        m = _scene_meshes(bodies)
        _lib.call("abd_body_boxes", _stream(), m.c, _lib.ptr(m.xk), float(dialation),
                  _lib.ptr(bvh.lowers), _lib.ptr(bvh.uppers))
        return bvh

    def body_bvh_to_traj(self, bodies, bvh_body: BoxSet, bvh_traj: BoxSet) -> BoxSet:
        m = _scene_meshes(bodies)
        bvh_traj.lowers.copy_(bvh_body.lowers)
        bvh_traj.uppers.copy_(bvh_body.uppers)
        _lib.call("abd_body_boxes_traj", _stream(), m.c, _lib.ptr(m.xk), _lib.ptr(bvh_traj.lowers),
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


def cull(ij_list, _bvh_list1: Boxes, _bvh_list2: Optional[Boxes] = None) -> torch.Tensor:
    """culling.py:8 / :32-52 (cull_ semantics, defect D1 fixed).  With a second box
    set: PT candidates (points of set 1 vs triangles of set 2, both body orders);
    without: EE candidates."""
    s1 = _scene_boxes(_bvh_list1)
    if _bvh_list2 is None:
        return _cull_prims(_lib.CULL_EE, ij_list, s1, s1)
    return _cull_prims(_lib.CULL_PT, ij_list, s1, _scene_boxes(_bvh_list2))


def cull_vg(lowers: torch.Tensor, _bvh_list: Boxes, bodies, dhat: Optional[float] = None,
            ground: Optional[float] = None) -> torch.Tensor:
    """culling.py:54-76: vertex-ground candidates (body, vertex).  dhat and ground
    default to the reference's module constants (const_params.py: dhat 1e-3) and the
    scene's floor height."""
    m = _scene_meshes(bodies)
    _bvh_list = _scene_boxes(_bvh_list)
    if dhat is None:
        dhat = ContactParams().dhat
    if ground is None:
        ground = float(m.spec.ground)
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
