"""Contact/solver parameters (const_params.py:24-39 in the reference).

The reference kernels read module constants (dt, dhat, kappa, stiffness, ...)
while `AffineBodySimulator.simulator_args` (abd.py:108-121) defines different
values that nothing reads.  Here every value is a runtime field of
`ContactParams`, passed through the C ABI as `AbdParams`.  The defaults are the
benchmark values: dhat 1e-3 (d2hat 1e-6), kappa 1e7, orthogonality stiffness
1e8, Newton tolerance 1e-4 on g.g, dt 0.01, max_iter 10, c1 1e-4, alpha_min
1e-3, gravity -9.8, mass 1e3, I0 1e2.
"""
from __future__ import annotations

import ctypes
import dataclasses


class AbdParamsC(ctypes.Structure):
    """Mirror of `AbdParams` in include/abd_hip.h (field order matters)."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "dt", "dhat", "d2hat", "kappa", "stiffness", "gravity", "ground", "mass", "I0", "c1", "alpha_min", "tol",
        "solver_tol")] + [(n, ctypes.c_int32) for n in (
            "max_iter", "solver_maxiter", "pt", "ee", "vg", "solver", "psd")]


SOLVER_BICGSTAB = 0
SOLVER_CG_JACOBI = 1
SOLVER_CG_BLOCK = 2    # CG with the 12x12 block-Jacobi preconditioner (beyond the reference)
SOLVER_LDLT = 3        # direct block LDL^T, block-Jacobi CG above 64 bodies (beyond the reference)


@dataclasses.dataclass
class ContactParams:
    dt: float = 0.01
    dhat: float = 1e-3
    kappa: float = 1e7
    stiffness: float = 1e8
    gravity: float = -9.8
    ground: float = -0.5
    mass: float = 1e3
    I0: float = 1e2
    c1: float = 1e-4
    alpha_min: float = 1e-3
    tol: float = 1e-4
    max_iter: int = 10
    solver_tol: float = 1e-4       # abd.py:244 bicgstab(hess, g, dq, 1e-4)
    solver_maxiter: int = 0        # 0 -> system size
    pt: bool = True                # const_params.py:42-44 contact-type switches
    ee: bool = True
    vg: bool = True
    solver: int = SOLVER_BICGSTAB
    # 0: the reference's contact Hessian (indefinite); 1: contact-space PSD projection;
    # 2: exact Hessian of the barrier energy; 3: full IPC (closest feature pairs, edge-edge
    # mollifier, per-contact 12x12 PSD projection) — 1, 2 and 3 go beyond the reference
    psd: int = 0

    @property
    def d2hat(self) -> float:
        return self.dhat * self.dhat

    def to_c(self) -> AbdParamsC:
        if int(self.psd) not in (0, 1, 2, 3):
            raise ValueError(f"psd must be 0, 1, 2 or 3, not {self.psd}")
        return AbdParamsC(self.dt, self.dhat, self.d2hat, self.kappa, self.stiffness, self.gravity, self.ground,
                          self.mass, self.I0, self.c1, self.alpha_min, self.tol, self.solver_tol, int(self.max_iter),
                          int(self.solver_maxiter), int(bool(self.pt)), int(bool(self.ee)), int(bool(self.vg)),
                          int(self.solver), int(self.psd))
