"""Continuous collision detection (ccd.py / cubic_roots.py in the reference).

`toi(meshes, pt_list, ee_list, vg_list)` is the device part of
AffineBodySimulator.ccd (abd.py:284-299, D4: the in-repo PT/EE cubic TOI on the
correct lists) over trajectories meshes.x -> meshes.xk.  `cubic_roots` and
`collision_time` are the batched forms of cubic_roots.py:95-193 and
ccd.py:133-186.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .const_params import ContactParams


def toi_device(meshes, params: ContactParams, pt_list, ee_list, vg_list, out=None, x_t0=None, x_t1=None):
    out = out if out is not None else torch.ones(1, dtype=torch.float64, device=meshes.device)
    x0 = meshes.x if x_t0 is None else x_t0
    x1 = meshes.xk if x_t1 is None else x_t1
    pc = params.to_c()
    _lib.call("abd_toi", _lib.stream_handle(), ctypes.byref(meshes.c), ctypes.byref(pc), _lib.ptr(x0), _lib.ptr(x1),
              _lib.ptr(pt_list), int(pt_list.shape[0]), _lib.ptr(ee_list), int(ee_list.shape[0]), _lib.ptr(vg_list),
              int(vg_list.shape[0]), _lib.ptr(out))
    return out


def toi(meshes, params: ContactParams, pt_list, ee_list, vg_list) -> float:
    """min(1, TOIs < 1) over every candidate (toi_vg + toi_ee + toi_pt)."""
    return float(toi_device(meshes, params, pt_list, ee_list, vg_list).item())


def cubic_roots(coef: torch.Tensor, x0: float, x1: float):
    """coef [n,4] constant-first -> (n_roots [n] int32, roots [n,3])."""
    n = int(coef.shape[0])
    coef = coef.contiguous()  # held for the duration of the launch
    nr = torch.zeros(n, dtype=torch.int32, device=coef.device)
    roots = torch.zeros((n, 3), dtype=torch.float64, device=coef.device)
    _lib.call("abd_cubic_roots", _lib.stream_handle(), _lib.ptr(coef), n, float(x0), float(x1),
              _lib.ptr(nr), _lib.ptr(roots))
    return nr, roots


def collision_time(kind: int, p_t0: torch.Tensor, p_t1: torch.Tensor) -> torch.Tensor:
    """kind 0 pt_collision_time, 1 ee_collision_time; p_t0/p_t1 [n,4,3] -> t [n]."""
    n = int(p_t0.shape[0])
    p_t0, p_t1 = p_t0.contiguous(), p_t1.contiguous()  # held for the duration of the launch
    t = torch.zeros(n, dtype=torch.float64, device=p_t0.device)
    _lib.call("abd_collision_time", _lib.stream_handle(), int(kind), _lib.ptr(p_t0), _lib.ptr(p_t1), n,
              _lib.ptr(t))
    return t
