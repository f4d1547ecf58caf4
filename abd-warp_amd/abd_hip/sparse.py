"""Block-sparse Newton matrix and its solvers (sparse.py / bsr_utils.py and
warp.sparse / warp.optim.linear as used at abd.py:240-244).

`BsrMatrix` couples the body-pair structure (row_ptr / cols / slots over 12x12
body blocks) with the slot values (`blocks`, [(Nb+2Nij)*16,3,3]).
`bicgstab(A, b, x, tol)` and `bsr_cg(A, x, b, ...)` solve in one workgroup
launch, x warm-started in place.
"""
from __future__ import annotations

import dataclasses

import torch

from . import _lib


@dataclasses.dataclass
class BsrMatrix:
    nb: int
    row_ptr: torch.Tensor
    cols: torch.Tensor
    slots: torch.Tensor
    blocks: torch.Tensor

    @property
    def shape(self):
        return (12 * self.nb, 12 * self.nb)


def bsr_from_pairs(nb: int, pairs, blocks: torch.Tensor) -> BsrMatrix:
    """bsr_set_from_triplets (abd.py:243) for the slot layout of `pairs`."""
    return BsrMatrix(nb, pairs.row_ptr, pairs.cols, pairs.slots, blocks)


class SolveInfo:
    def __init__(self, device):
        self.iters = torch.zeros(1, dtype=torch.int32, device=device)
        self.resid = torch.zeros(1, dtype=torch.float64, device=device)


def _solve(method, A: BsrMatrix, b, x, tol, maxiter, check_every, info):
    info = info or SolveInfo(b.device)
    _lib.call("abd_bsr_solve", _lib.stream_handle(), int(method), A.nb, _lib.ptr(A.row_ptr), _lib.ptr(A.cols),
              _lib.ptr(A.slots), _lib.ptr(A.blocks), _lib.ptr(b), _lib.ptr(x), float(tol), int(maxiter),
              int(check_every), _lib.ptr(info.iters), _lib.ptr(info.resid))
    return info


def bicgstab(A: BsrMatrix, b, x, tol=1e-4, maxiter=0, check_every=10, info=None):
    """warp.optim.linear.bicgstab(A, b, x, tol) as called at abd.py:244 (relative tol)."""
    return _solve(_lib.SOLVER_BICGSTAB, A, b, x, tol, maxiter, check_every, info)


def bsr_cg(A: BsrMatrix, x, b, max_iters=0, tol=1e-4, check_every=10, info=None, block_precond=False):
    """bsr_utils.py:126-235 (diagonal preconditioner; absolute tol^2 n on r.z).  block_precond
    (beyond the reference) preconditions with each body's 12x12 diagonal block instead."""
    method = _lib.SOLVER_CG_BLOCK if block_precond else _lib.SOLVER_CG_JACOBI
    return _solve(method, A, b, x, tol, max_iters, check_every, info)


def bsr_ldlt(A: BsrMatrix, x, b, max_iters=0, tol=1e-4, check_every=10, info=None):
    """Direct block LDL^T of the 12x12 body blocks (beyond the reference) for systems of at
    most 64 bodies; larger systems, or one that is not positive definite as factored, take
    bsr_cg (12x12 block-Jacobi) with max_iters / tol / check_every.  info.iters = 1 for a direct solve."""
    return _solve(_lib.SOLVER_LDLT, A, b, x, tol, max_iters, check_every, info)
