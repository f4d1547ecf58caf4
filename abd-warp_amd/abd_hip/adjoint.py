"""Gradients of a simulated trajectory by the implicit-function theorem.

Each frame of AffineBodySimulator.step (abd.py:212-276) returns the minimiser
q* of the frame energy E(q; q~) = inertia(q - q~) + orthogonality(q) +
barrier(x(q)), with the inertial target q~ = q0 + dt q0' + dt^2 (0, g, 0)
(orthogonal_energy.py:129-135), followed by q0' <- (q* - q0) / dt, q0 <- q*
(abd.py:355-362).  At the minimiser grad E = 0, so dq*/dq~ = H^-1 M, with H
the exact Hessian of E at q* (psd = 2: the barrier's exact second derivative,
which the reference's Newton matrix approximates) and M the diagonal mass
matrix (m for p, I0 for the rows of A).  One backward step is one extra solve
against that Hessian:

    lam    = dL/dq_{n+1} + dL/dq'_{n+1} / dt
    qt_bar = M H^-1 lam
    dL/dq_n  = qt_bar - dL/dq'_{n+1} / dt,      dL/dq'_n = dt qt_bar.

The forward pass keeps one checkpoint per frame (A0, p0, A0', p0': 192 B per
body); the backward pass rebuilds H at each checkpointed q* on the GPU
(abd_sim_hessian_solve) and never re-runs a forward frame.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import torch

from . import _lib
from .abd import AffineBodySimulator
from .const_params import SOLVER_CG_JACOBI, ContactParams
from .scenes import SceneSpec


def _to_q(A: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
    """(A [nb,3,3], p [nb,3]) -> the Newton layout [4 nb, 3]: per body p, a1, a2, a3."""
    return torch.cat([p[:, None, :], A], dim=1).reshape(-1, 3)


def _from_q(q: torch.Tensor, nb: int) -> Tuple[torch.Tensor, torch.Tensor]:
    v = q.reshape(nb, 4, 3)
    return v[:, 1:, :], v[:, 0, :]


class Rollout:
    """A checkpointed forward simulation and its adjoint (one scene).

    `params.psd` must be 2 (exact Hessian) and the Newton loop should be run to
    a tight tolerance (`params.tol`, `params.max_iter`), since the adjoint is
    exact for the minimiser of each frame."""

    def __init__(self, spec: SceneSpec, params: ContactParams, device="cuda", solve_tol: float = 1e-12,
                 solve_method: int = SOLVER_CG_JACOBI, solve_maxiter: int = 20000):
        if params.psd != 2:
            raise ValueError("the adjoint needs the exact Hessian: ContactParams(psd=2)")
        if spec.n_scenes > 1:
            raise ValueError("one scene per Rollout")
        self.spec = spec
        self.params = params
        self.sim = AffineBodySimulator(spec, params, device, native=True)
        self.nb = spec.n_bodies
        self.device = self.sim.device
        st = self.sim.states
        self.mass_q = torch.stack([st.mass, st.I0, st.I0, st.I0], dim=1).reshape(-1, 1)  # [4 nb, 1]
        # the Hessian of a frame with vertex-ground contacts only is SPD: Jacobi CG; BiCGSTAB
        # (solve_method = 0) for scenes whose exact contact Hessian is indefinite
        self.solve_tol = solve_tol
        self.solve_method = solve_method
        self.solve_maxiter = solve_maxiter
        self.checkpoints: List[torch.Tensor] = []
        self.solver_iters: List[int] = []
        self.newton_iters: List[int] = []

    # ---------------------------------------------------------- forward
    def _save(self):
        st = self.sim.states
        self.checkpoints.append(torch.cat([st.A0.reshape(self.nb, 9), st.p0, st.Adot.reshape(self.nb, 9), st.pdot],
                                          dim=1).clone())

    def forward(self, pdot0: Optional[torch.Tensor], steps: int) -> Tuple[torch.Tensor, torch.Tensor]:
        """Runs `steps` frames from the scene's initial state (initial velocities
        replaced by pdot0 when given); returns the final (A, p)."""
        self.sim.reset()
        if pdot0 is not None:
            self.sim.states.pdot.copy_(pdot0.detach().to(torch.float64))
        self.checkpoints = []
        self.newton_iters = []
        self._save()
        for f in range(steps):
            self.newton_iters.append(self.sim.step(f))
            self._save()
        st = self.sim.states
        return st.A0.clone(), st.p0.clone()

    # --------------------------------------------------------- backward
    def backward(self, gA: torch.Tensor, gp: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor,
                                                                     torch.Tensor]:
        """dL/d(final A, p) -> dL/d(initial A0, p0, A0', p0')."""
        nb, dt = self.nb, self.params.dt
        st = self.sim.states
        gq = _to_q(gA.to(torch.float64), gp.to(torch.float64)).contiguous()
        gv = torch.zeros_like(gq)
        y = torch.zeros_like(gq)
        it = ctypes.c_int32(0)
        res = ctypes.c_double(0.0)
        self.solver_iters = []
        stream = _lib.stream_handle()
        for n in range(len(self.checkpoints) - 2, -1, -1):
            ck = self.checkpoints[n + 1]
            st.A.copy_(ck[:, :9].reshape(nb, 3, 3))
            st.p.copy_(ck[:, 9:12])
            lam = (gq + gv / dt).contiguous()
            _lib.call("abd_sim_hessian_solve", self.sim._sim, stream, self.solve_method, self.solve_tol,
                      self.solve_maxiter, _lib.ptr(lam), _lib.ptr(y), ctypes.byref(it), ctypes.byref(res))
            self.solver_iters.append(int(it.value))
            qt_bar = self.mass_q * y
            gq = qt_bar - gv / dt
            gv = dt * qt_bar
        gA0, gp0 = _from_q(gq, nb)
        gAdot, gpdot = _from_q(gv, nb)
        return gA0, gp0, gAdot, gpdot


class RolloutFunction(torch.autograd.Function):
    """final (A, p) = rollout(initial velocities pdot0) with the IFT adjoint as backward."""

    @staticmethod
    def forward(ctx, pdot0: torch.Tensor, rollout: Rollout, steps: int):
        A, p = rollout.forward(pdot0, steps)
        ctx.rollout = rollout
        return A, p

    @staticmethod
    def backward(ctx, gA, gp):
        nb = ctx.rollout.nb
        gA = torch.zeros((nb, 3, 3), dtype=torch.float64, device=ctx.rollout.device) if gA is None else gA
        gp = torch.zeros((nb, 3), dtype=torch.float64, device=ctx.rollout.device) if gp is None else gp
        _, _, _, gpdot = ctx.rollout.backward(gA, gp)
        return gpdot, None, None


def centre_of_mass(A: torch.Tensor, p: torch.Tensor, rest_centroid: torch.Tensor) -> torch.Tensor:
    """Vertex centroid of each body: p + A^T c (c the rest-space vertex mean)."""
    return p + (A * rest_centroid[:, :, None]).sum(dim=1)


def rest_centroids(spec: SceneSpec, device="cuda") -> torch.Tensor:
    import numpy as np
    c = np.stack([spec.x0[spec.voff[b]:spec.voff[b + 1]].mean(0) for b in range(spec.n_bodies)])
    return torch.as_tensor(c, dtype=torch.float64, device=device)
