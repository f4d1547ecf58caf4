"""IPC contact energy, gradient and Hessian (ipc.py:15-62 in the reference).

`IPCContactEnergy.energy([ee_list, pt_list, vg_list, meshes])` and
`IPCContactEnergy.gh([pt_list, ee_list, vg_list, nij, meshes, g, blocks])`
keep the reference's input lists.  `meshes` is the scene (`Meshes`); `g`
([4Nb,3] float64) and `blocks` ([(Nb+2Nij)*16,3,3] float64, the slot layout of
ipcsk/put.py:26-49) are caller-owned and accumulated into.  Both run
entirely in libabd_hip.so.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .const_params import ContactParams


class IPCContactEnergy:
    def __init__(self, params: ContactParams):
        self.params = params
        self._E = None

    def energy(self, inputs):
        """ipc.py:19-46 (D3: each list goes to its own formula).  Returns a float
        (the reference's `E.numpy()[0]`, a host read-back)."""
        return float(self.energy_device(inputs).item())

    def energy_device(self, inputs) -> torch.Tensor:
        ee_list, pt_list, vg_list, meshes = inputs[:4]
        if self._E is None or self._E.device != meshes.device:
            self._E = torch.zeros(1, dtype=torch.float64, device=meshes.device)
        pc = self.params.to_c()
        _lib.call("abd_contact_energy", _lib.stream_handle(), ctypes.byref(meshes.c), ctypes.byref(pc),
                  _lib.ptr(meshes.xk), _lib.ptr(pt_list), int(pt_list.shape[0]), _lib.ptr(ee_list),
                  int(ee_list.shape[0]), _lib.ptr(vg_list), int(vg_list.shape[0]), _lib.ptr(self._E))
        return self._E

    def gh(self, inputs, return_active: bool = False):
        """ipc.py:54-62: adds the barrier gradient into g and the PSD-projected
        Hessian into blocks, at meshes.x.  `nij` may be the int the reference
        passes (then the body pairs of the scene's latest ij_list() are used) or
        the BodyPairs object itself.  Returns the per-body highlight of the
        reference (bodies J of active EE contacts, ipc.py:59-60)."""
        pt_list, ee_list, vg_list, nij, meshes, g, blocks = inputs[:7]
        pairs = nij if hasattr(nij, "row_ptr") else meshes.pairs
        if pairs is None or (not hasattr(nij, "row_ptr") and pairs.nij != int(nij)):
            raise ValueError("gh(): no body-pair structure for nij; call ij_list() on this scene first")
        nb = meshes.n_bodies
        hl = torch.zeros(nb, dtype=torch.uint8, device=meshes.device)
        nact = ctypes.c_int64(0)
        pc = self.params.to_c()
        _lib.call("abd_contact_gh", _lib.stream_handle(), ctypes.byref(meshes.c), ctypes.byref(pc),
                  _lib.ptr(meshes.x), _lib.ptr(pt_list), int(pt_list.shape[0]), _lib.ptr(ee_list),
                  int(ee_list.shape[0]), _lib.ptr(vg_list), int(vg_list.shape[0]), pairs.nij,
                  _lib.ptr(pairs.row_ptr), _lib.ptr(pairs.slots), _lib.ptr(g), _lib.ptr(blocks), _lib.ptr(hl),
                  ctypes.byref(nact))
        self.n_active = nact.value
        self.highlight_bits = hl
        highlight = (hl & 2) != 0
        return (highlight, nact.value) if return_active else highlight


def pair_gh(kind: int, params: ContactParams, x4: torch.Tensor, r4: torch.Tensor):
    """Per-contact gh (ipcsk/pt_ipc.py:111-185, ipcsk/ee_ipc.py:94-143) for n contacts:
    x4/r4 [n,4,3] current/rest positions -> (active [n] bool, g24 [n,24], H24 [n,24,24])."""
    n = int(x4.shape[0])
    dev = x4.device
    x4, r4 = x4.contiguous(), r4.contiguous()  # held for the duration of the launch
    act = torch.zeros(n, dtype=torch.uint8, device=dev)
    g24 = torch.zeros((n, 24), dtype=torch.float64, device=dev)
    H24 = torch.zeros((n, 24, 24), dtype=torch.float64, device=dev)
    pc = params.to_c()
    _lib.call("abd_pair_gh", _lib.stream_handle(), int(kind), ctypes.byref(pc), _lib.ptr(x4), _lib.ptr(r4), n,
              _lib.ptr(act), _lib.ptr(g24), _lib.ptr(H24))
    return act.bool(), g24, H24
