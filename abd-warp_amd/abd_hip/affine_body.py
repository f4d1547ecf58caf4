"""Affine-body state and mesh data on the GPU (affine_body.py in the reference).

`AffineBodyStates` keeps the reference's field names (affine_body.py:8-22) as
torch float64 tensors: A, Ak, A0, Adot [Nb,3,3] (row r = a_{r+1}; a vertex maps
to x = A^T x0 + p, abd.py:313-316), p, pk, p0, pdot [Nb,3], mass, I0 [Nb].

The reference keeps one `WarpMesh` struct per body (affine_body.py:85-95) and a
`wp.array` of them.  `Meshes` replaces that array with one flattened scene:
x0 / x / xk / x_view [V,3] for all bodies, triangles [F,3] / edges [E,2] in
body-local vertex ids, and per-body offsets.  `meshes[b]` returns a `WarpMesh`
whose fields are views into the flattened arrays, so per-body code reads the
same fields as before.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np
import torch

from . import _lib
from .scenes import SceneSpec


@dataclasses.dataclass
class AffineBodyStates:
    A: torch.Tensor
    p: torch.Tensor
    Ak: torch.Tensor
    pk: torch.Tensor
    Adot: torch.Tensor
    pdot: torch.Tensor
    A0: torch.Tensor
    p0: torch.Tensor
    mass: Optional[torch.Tensor] = None
    I0: Optional[torch.Tensor] = None


def affine_body_states_empty(n_bodies: int, device="cuda") -> AffineBodyStates:
    """affine_body.py:24-40 (mass and I0 are allocated here: D13)."""
    z33 = lambda: torch.zeros((n_bodies, 3, 3), dtype=torch.float64, device=device)
    z3 = lambda: torch.zeros((n_bodies, 3), dtype=torch.float64, device=device)
    return AffineBodyStates(A=z33(), p=z3(), Ak=z33(), pk=z3(), Adot=z33(), pdot=z3(), A0=z33(), p0=z3(),
                            mass=torch.zeros(n_bodies, dtype=torch.float64, device=device),
                            I0=torch.zeros(n_bodies, dtype=torch.float64, device=device))


@dataclasses.dataclass
class WarpMesh:
    """Per-body view (affine_body.py:85-95): every field aliases the scene arrays."""
    id: int
    x0: torch.Tensor
    x: torch.Tensor
    xk: torch.Tensor
    x_view: torch.Tensor
    triangles: torch.Tensor
    edges: torch.Tensor
    meshes: Optional["Meshes"] = None  # the scene the views belong to


class Meshes:
    """The scene's meshes, flattened on the device (one object for all bodies)."""

    def __init__(self, spec: SceneSpec, device="cuda"):
        self.spec = spec
        self.device = torch.device(device)
        dev = self.device
        nb = spec.n_bodies
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)
        f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        self.n_bodies = nb
        self.voff = i32(spec.voff)
        self.foff = i32(spec.foff)
        self.eoff = i32(spec.eoff)
        vcount = np.diff(spec.voff)
        fcount = np.diff(spec.foff)
        ecount = np.diff(spec.eoff)
        self.vert_body = i32(np.repeat(np.arange(nb, dtype=np.int32), vcount))
        self.tri_body = i32(np.repeat(np.arange(nb, dtype=np.int32), fcount))
        self.edge_body = i32(np.repeat(np.arange(nb, dtype=np.int32), ecount))
        self.triangles = i32(spec.tris.reshape(-1, 3))
        self.edges = i32(spec.edges.reshape(-1, 2))
        self.x0 = f64(spec.x0.reshape(-1, 3))
        self.x = self.x0.clone()
        self.xk = self.x0.clone()
        self.x_view = self.x0.clone()
        self.body_scene = i32(spec.body_scene)
        self.mass = f64(spec.mass)
        self.I0 = f64(spec.I0)
        # independent scenes: contiguous body ranges and per-scene pair offsets
        bs = np.asarray(spec.body_scene, np.int64)
        if nb and np.any(np.diff(bs) < 0):
            raise ValueError("bodies of a scene must be contiguous and scenes in increasing order")
        S = int(bs.max()) + 1 if nb else 0
        sb = np.searchsorted(bs, np.arange(S + 1)).astype(np.int32)
        nbs = np.diff(sb).astype(np.int64)
        sp = np.concatenate([[0], np.cumsum(nbs * (nbs - 1) // 2)]).astype(np.int64)
        self.n_scenes = S
        self.scene_boff_host = sb
        self.scene_boff = i32(sb)
        self.scene_poff = torch.as_tensor(sp, device=dev)
        self.max_verts = int(vcount.max()) if nb else 0
        self.max_tris = int(fcount.max()) if nb else 0
        self.max_edges = int(ecount.max()) if nb else 0
        self.n_verts, self.n_tris, self.n_edges = spec.n_verts, spec.n_tris, spec.n_edges
        self.c = _lib.AbdScene(
            nb, spec.n_verts, spec.n_tris, spec.n_edges, self.max_verts, self.max_tris, self.max_edges,
            _lib.ptr(self.voff), _lib.ptr(self.foff), _lib.ptr(self.eoff), _lib.ptr(self.vert_body),
            _lib.ptr(self.tri_body), _lib.ptr(self.edge_body), _lib.ptr(self.triangles), _lib.ptr(self.edges),
            _lib.ptr(self.x0), _lib.ptr(self.body_scene), _lib.ptr(self.mass), _lib.ptr(self.I0),
            S, _lib.ptr(self.scene_boff), _lib.ptr(self.scene_poff), int(sp[-1]))
        # body pairs of the most recent ij_list(), used by gh()'s slot layout
        self.pairs = None

    # the reference indexes the WarpMesh array per body
    def __len__(self) -> int:
        return self.n_bodies

    def __getitem__(self, b: int) -> WarpMesh:
        if not 0 <= b < self.n_bodies:
            raise IndexError(b)
        v0, v1 = int(self.spec.voff[b]), int(self.spec.voff[b + 1])
        f0, f1 = int(self.spec.foff[b]), int(self.spec.foff[b + 1])
        e0, e1 = int(self.spec.eoff[b]), int(self.spec.eoff[b + 1])
        return WarpMesh(b, self.x0[v0:v1], self.x[v0:v1], self.xk[v0:v1], self.x_view[v0:v1],
                        self.triangles[f0:f1], self.edges[e0:e1], self)

    @property
    def shape(self):
        return (self.n_bodies,)

    def states(self) -> AffineBodyStates:
        """Initial states of the scene (abd.py:33-38 gather + q_gets_q0)."""
        st = affine_body_states_empty(self.n_bodies, self.device)
        sp = self.spec
        for name, val in (("A0", sp.A), ("p0", sp.p), ("Adot", sp.Adot), ("pdot", sp.pdot)):
            getattr(st, name).copy_(torch.as_tensor(np.asarray(val, np.float64)))
        st.A.copy_(st.A0)
        st.p.copy_(st.p0)
        st.Ak.copy_(st.A0)
        st.pk.copy_(st.p0)
        st.mass.copy_(self.mass)
        st.I0.copy_(self.I0)
        return st
