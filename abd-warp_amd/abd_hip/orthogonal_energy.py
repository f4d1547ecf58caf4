"""Inertia + orthogonality energy and the Newton matrix triplets
(orthogonal_energy.py in the reference).

`InertialEnergy(params, meshes).energy(states)` / `.gradient(g, states)` /
`.hessian(blocks, states)` keep the reference's call shapes
(orthogonal_energy.py:7-22): energy at (Ak, pk), gradient and Hessian at
(A, p); gradient overwrites g, Hessian adds into the diagonal slots.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .const_params import ContactParams


class InertialEnergy:
    def __init__(self, params: ContactParams, meshes):
        self.params = params
        self.meshes = meshes
        self.e = torch.zeros(1, dtype=torch.float64, device=meshes.device)

    def energy_device(self, states) -> torch.Tensor:
        pc = self.params.to_c()
        _lib.call("abd_inertia_energy", _lib.stream_handle(), ctypes.byref(self.meshes.c), ctypes.byref(pc),
                  _lib.ptr(states.Ak), _lib.ptr(states.pk), _lib.ptr(states.A0), _lib.ptr(states.p0),
                  _lib.ptr(states.Adot), _lib.ptr(states.pdot), _lib.ptr(self.e))
        return self.e

    def energy(self, states) -> float:
        return float(self.energy_device(states).item())

    def gradient(self, g, states):
        pc = self.params.to_c()
        _lib.call("abd_inertia_gradient", _lib.stream_handle(), ctypes.byref(self.meshes.c), ctypes.byref(pc),
                  _lib.ptr(states.A), _lib.ptr(states.p), _lib.ptr(states.A0), _lib.ptr(states.p0),
                  _lib.ptr(states.Adot), _lib.ptr(states.pdot), _lib.ptr(g))

    def hessian(self, blocks, states):
        pc = self.params.to_c()
        _lib.call("abd_inertia_hessian", _lib.stream_handle(), ctypes.byref(self.meshes.c), ctypes.byref(pc),
                  _lib.ptr(states.A), _lib.ptr(blocks))


def set_triplets(n_bodies: int, n_ij: int, ij_list: torch.Tensor, rows: torch.Tensor, cols: torch.Tensor):
    """orthogonal_energy.py:138-165 _set_triplets over every slot (D5: off-diagonal
    slots included): rows/cols [16 (Nb + 2 Nij)] int32."""
    _lib.call("abd_set_triplets", _lib.stream_handle(), int(n_bodies), _lib.ptr(ij_list), int(n_ij), _lib.ptr(rows),
              _lib.ptr(cols))
