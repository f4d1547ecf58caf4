"""The Newton driver that calls the contact path (abd.py in the reference).

`AffineBodySimulator` mirrors AffineBodySimulator (abd.py:16-362) method by
method — q_gets_q0, compute_energy, dot, line_search, ij_list,
trajectory_intersection_set, collision_set, V_gets_V, step, ccd,
update_mesh_vertex, update_q / update_qk / update_q0qdot — with every stage
running in libabd_hip.so.  Differences from the reference, all deliberate:
  D1  culling uses the BVH semantics of cull_ (culling.py:32-52), not the stub;
  D3  compute_energy passes each list to its own formula (abd.py:137);
  D4  ccd uses the in-repo cubic TOI on the correct lists (abd.py:282-299);
  D5  the Newton matrix has Nb + 2 Nij slots with full off-diagonal coupling;
  D7  parameters come from ContactParams, not module constants;
  D12 every vertex update uses x = A^T x0 + p.
The scene is a SceneSpec (procedural meshes) instead of config.json + bgeo assets.
"""
from __future__ import annotations

import ctypes

from typing import Optional

import torch

from . import _lib
from .affine_body import Meshes
from .ccd import toi_device
from .const_params import ContactParams
from .culling import BodyPairs, BvhBuilder, cull, cull_vg, ij_list as _ij_list
from .ipc import IPCContactEnergy
from .orthogonal_energy import InertialEnergy
from .scenes import SceneSpec
from .sparse import SolveInfo, bicgstab, bsr_cg, bsr_from_pairs, bsr_ldlt


class AffineBodySimulator:
    """native=True runs step() in the library's C++ frame driver (abd_sim_step): the
    same kernels in the same order, without a Python round trip per stage.

    A SceneSpec with several independent scenes (body_scene) needs native=True: the
    driver then runs each scene's own Newton loop, with one launch per stage for all
    scenes; `scene_iters()` gives each scene's Newton iterations of the last frame."""

    def __init__(self, spec: SceneSpec, params: Optional[ContactParams] = None, device="cuda", native=False):
        self.params = params if params is not None else ContactParams(ground=spec.ground)
        self.meshes = Meshes(spec, device)
        self.warp_meshes = self.meshes
        self.device = self.meshes.device
        nb = spec.n_bodies
        self.n_bodies = nb
        self.states = self.meshes.states()
        self.g = torch.zeros((4 * nb, 3), dtype=torch.float64, device=self.device)
        self.dq = torch.zeros_like(self.g)
        self._blocks = torch.zeros((16 * nb, 3, 3), dtype=torch.float64, device=self.device)
        self.inertia = InertialEnergy(self.params, self.meshes)
        self.ipc_contact = IPCContactEnergy(self.params)
        self.bb = BvhBuilder(self.meshes)
        m = self.meshes
        self.update_mesh_vertex("x")
        self.update_mesh_vertex("xk")
        dh = self.params.dhat * 0.5
        self.bvh_bodies = self.bb.build_body_bvh(m, dh)
        self.bvh_body_traj = self.bb.build_body_bvh(m, dh)
        self.bvh_triangles = self.bb.bulid_triangle_bvh(m.x, m.triangles, 0.0)
        self.bvh_edges = self.bb.build_edge_bvh(m.x, m.edges, dh)
        self.bvh_points = self.bb.build_point_bvh(m.x, self.params.dhat)
        self.bvh_triangle_traj = self.bb.bulid_triangle_bvh(m.x, m.triangles, 0.0)
        self.bvh_edge_traj = self.bb.build_edge_bvh(m.x, m.edges, dh)
        self.bvh_point_traj = self.bb.build_point_bvh(m.x, self.params.dhat)
        self.highlight = torch.zeros(nb, dtype=torch.bool, device=self.device)
        self._dot = torch.zeros(2, dtype=torch.float64, device=self.device)
        self._toi = torch.ones(1, dtype=torch.float64, device=self.device)
        self.solve_info = SolveInfo(self.device)
        self.stats = dict(newton_iters=0, gh_pt=0, gh_ee=0, gh_vg=0, gh_active=0, solver_iters=0, energy_evals=0,
                          ccd_pt=0, ccd_ee=0, frames=0)
        self.profile_gh = False
        self.gh_time_ms = 0.0
        self._gh_events = []
        self._sim = None
        if self.meshes.n_scenes > 1 and not native:
            raise ValueError("independent scenes run in the native driver: pass native=True")
        if native:
            self._make_native()

    def _make_native(self):
        import ctypes
        st, m = self.states, self.meshes
        self._state_c = _lib.AbdSimState(*[_lib.ptr(t) for t in (st.A, st.p, st.Ak, st.pk, st.A0, st.p0, st.Adot,
                                                                  st.pdot, m.x, m.xk, m.x_view, self.g, self.dq)])
        self._params_c = self.params.to_c()
        h = ctypes.c_void_p(0)
        _lib.call("abd_sim_create", m.c, self._params_c, ctypes.byref(self._state_c), ctypes.byref(h))
        self._sim = h
        self._native_stats = (ctypes.c_int64 * 10)()

    @classmethod
    def from_config(cls, config_file: str = "config.json", device="cuda", native=False) -> "AffineBodySimulator":
        """abd.py:21 AffineBodySimulator(config_file): scene and parameters from the
        reference's config.json / scenes/*.json schema (abd_hip.scene_io)."""
        from .scene_io import load_config
        cfg = load_config(config_file)
        sim = cls(cfg.scene, cfg.params, device, native=native)
        sim.frame_in, sim.frame_out = cfg.frame_in, cfg.frame_out
        return sim

    def autotune(self) -> dict:
        """Picks the culling launch shapes (abd_sim_autotune) at the current state; the
        results of every later frame are unchanged by the choice."""
        if self._sim is None:
            raise RuntimeError("autotune() needs the native driver")
        t = _lib.AbdTuning()
        _lib.call("abd_sim_autotune", self._sim, _lib.stream_handle(), ctypes.byref(t))
        return {"cull_chunk": t.cull_chunk, "cull_grid": t.cull_grid, "pair_grid": t.pair_grid}

    def set_graphs(self, on: bool):
        """Graph replay of the batched native driver's Newton iterations (abd_sim_set_graphs;
        on by default, results unchanged)."""
        if self._sim is None:
            raise RuntimeError("set_graphs() needs the native driver")
        _lib.call("abd_sim_set_graphs", self._sim, 1 if on else 0)

    def graph_counts(self) -> dict:
        """Batched Newton iterations replayed from a graph, recorded and launched eagerly."""
        if self._sim is None:
            raise RuntimeError("graph_counts() needs the native driver")
        out = (ctypes.c_int64 * 3)()
        _lib.call("abd_sim_graph_counts", self._sim, out)
        return {"replayed": out[0], "recorded": out[1], "eager": out[2]}

    def scene_iters(self) -> list:
        """Newton iterations of each scene in the last native frame."""
        import ctypes
        if self._sim is None:
            raise RuntimeError("scene_iters() needs the native driver")
        out = (ctypes.c_int32 * max(self.meshes.n_scenes, 1))()
        _lib.call("abd_sim_scene_iters", self._sim, out)
        return list(out)

    def __del__(self):
        if getattr(self, "_sim", None) is not None:
            try:
                _lib.call("abd_sim_destroy", self._sim)
            except Exception:
                pass
            self._sim = None

    def _step_native(self) -> int:
        before = list(self._native_stats)
        _lib.call("abd_sim_step", self._sim, _lib.stream_handle(), self._native_stats)
        s = list(self._native_stats)
        d = [a - b for a, b in zip(s, before)]
        for k, name in enumerate(("newton_iters", "gh_pt", "gh_ee", "gh_vg", "gh_active", "energy_evals",
                                  "solver_iters", "frames", "ccd_pt", "ccd_ee")):
            # library order: iters, pt, ee, vg, active, energy evals, solver iters, frames, ccd pt, ccd ee
            self.stats[name] += d[k]
        return d[0]

    # ------------------------------------------------------------ states
    def reset(self):
        """abd.py:97-104 (copies into the existing buffers, which the native driver holds)"""
        fresh = self.meshes.states()
        for name in ("A", "p", "Ak", "pk", "A0", "p0", "Adot", "pdot", "mass", "I0"):
            getattr(self.states, name).copy_(getattr(fresh, name))
        self.dq.zero_()
        self.update_mesh_vertex("x")

    def q_gets_q0(self):
        """abd.py:124-128"""
        st = self.states
        st.A.copy_(st.A0)
        st.p.copy_(st.p0)
        st.Ak.copy_(st.A0)
        st.pk.copy_(st.p0)

    def update_mesh_vertex(self, field: str = "x"):
        """abd.py:305-324 (on the GPU; no host round trip)."""
        st = self.states
        A, p = (st.A0, st.p0) if field == "x_view" else (st.Ak, st.pk) if field == "xk" else (st.A, st.p)
        _lib.call("abd_project_vertices", _lib.stream_handle(), self.meshes.c, _lib.ptr(A), _lib.ptr(p),
                  _lib.ptr(getattr(self.meshes, field)))

    def update_q(self, alpha: float):
        st = self.states
        _lib.call("abd_update_q", _lib.stream_handle(), self.n_bodies, _lib.ptr(st.A), _lib.ptr(st.p),
                  _lib.ptr(self.dq), float(alpha), _lib.ptr(st.A), _lib.ptr(st.p))

    def update_qk(self, alpha: float):
        st = self.states
        _lib.call("abd_update_q", _lib.stream_handle(), self.n_bodies, _lib.ptr(st.A), _lib.ptr(st.p),
                  _lib.ptr(self.dq), float(alpha), _lib.ptr(st.Ak), _lib.ptr(st.pk))

    def update_q0qdot(self):
        st = self.states
        _lib.call("abd_update_q0qdot", _lib.stream_handle(), self.n_bodies, float(self.params.dt), _lib.ptr(st.A),
                  _lib.ptr(st.p), _lib.ptr(st.A0), _lib.ptr(st.p0), _lib.ptr(st.Adot), _lib.ptr(st.pdot))

    def V_gets_V(self, states=None):
        self.update_mesh_vertex("x")

    # ------------------------------------------------------------ energies
    def dot(self, a, b, slot: int = 0) -> float:
        """abd.py:143-144 (array_inner)."""
        _lib.call("abd_dot", _lib.stream_handle(), _lib.ptr(a), _lib.ptr(b), int(a.numel()),
                  _lib.ptr(self._dot[slot:slot + 1]))
        return float(self._dot[slot].item())

    def compute_energy(self, alpha: float = 0.0) -> float:
        """abd.py:133-141 (D3: lists passed in their own order)."""
        self.update_qk(alpha)
        self.update_mesh_vertex("xk")
        ij, ee_list, pt_list, vg_list = self.collision_set()
        e_c = float(self.ipc_contact.energy_device([ee_list, pt_list, vg_list, self.meshes]).item())
        e_i = float(self.inertia.energy_device(self.states).item())
        self.stats["energy_evals"] += 1
        return e_c + e_i

    def line_search(self, alpha_cap: float, E0: float):
        """abd.py:146-160 (Armijo test as written, D8)."""
        qTg = self.dot(self.dq, self.g, 1)
        alpha = alpha_cap
        p = self.params
        while True:
            E1 = self.compute_energy(alpha)
            wolfe = E1 < E0 + p.c1 * alpha * qTg
            if wolfe or alpha < p.alpha_min:
                break
            alpha *= 0.5
        return alpha, E1

    # ------------------------------------------------------------ culling
    def ij_list(self, bvh_bodies) -> BodyPairs:
        """abd.py:163-168"""
        return _ij_list(bvh_bodies, self.meshes)

    def trajectory_intersection_set(self):
        """abd.py:170-185: boxes swept from the last collision_set to xk."""
        m = self.meshes
        self.bb.body_bvh_to_traj(m, self.bvh_bodies, self.bvh_body_traj)
        ij = self.ij_list(self.bvh_body_traj)
        self.bb.triangle_bvh_to_traj(m.xk, m.triangles, self.bvh_triangles, self.bvh_triangle_traj)
        self.bb.edge_bvh_to_traj(m.xk, m.edges, self.bvh_edges, self.bvh_edge_traj)
        self.bb.point_bvh_to_traj(m.xk, self.bvh_points, self.bvh_point_traj)
        ee_list = cull(ij, self.bvh_edge_traj)
        pt_list = cull(ij, self.bvh_point_traj, self.bvh_triangle_traj)
        vg_list = cull_vg(self.bvh_body_traj.lowers, self.bvh_point_traj, m, self.params.dhat, self.params.ground)
        return ij, pt_list, ee_list, vg_list

    def collision_set(self):
        """abd.py:187-207 (all on xk)."""
        m = self.meshes
        dh = self.params.dhat * 0.5
        self.bb.update_body_bvh(m, dh, self.bvh_bodies)
        ij = self.ij_list(self.bvh_bodies)
        self.bb.update_triangle_bvh(m.xk, m.triangles, 0.0, self.bvh_triangles)
        self.bb.update_edge_bvh(m.xk, m.edges, dh, self.bvh_edges)
        self.bb.update_point_bvh(m.xk, self.params.dhat, self.bvh_points)
        ee_list = cull(ij, self.bvh_edges)
        pt_list = cull(ij, self.bvh_points, self.bvh_triangles)
        vg_list = cull_vg(self.bvh_bodies.lowers, self.bvh_points, m, self.params.dhat, self.params.ground)
        return ij, ee_list, pt_list, vg_list

    # ------------------------------------------------------------ Newton
    def _blocks_for(self, nij: int) -> torch.Tensor:
        n = 16 * (self.n_bodies + 2 * nij)
        if self._blocks.shape[0] < n:
            self._blocks = torch.zeros((n + n // 2, 3, 3), dtype=torch.float64, device=self.device)
        b = self._blocks[:n]
        b.zero_()
        return b

    def step(self, frame: int = 1):
        """abd.py:212-276: one frame."""
        if self._sim is not None:
            return self._step_native()
        states = self.states
        p = self.params
        self.q_gets_q0()
        E0 = self.compute_energy(alpha=0.0)
        it = 0
        while True:
            self.V_gets_V(states)
            self.update_mesh_vertex("xk")
            pairs, ee_list, pt_list, vg_list = self.collision_set()
            nij = pairs.nij
            blocks = self._blocks_for(nij)
            self.g.zero_()
            self.inertia.gradient(self.g, states)
            self.inertia.hessian(blocks, states)
            if self.profile_gh:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            hl, nact = self.ipc_contact.gh([pt_list, ee_list, vg_list, pairs, self.meshes, self.g, blocks],
                                           return_active=True)
            if self.profile_gh:
                e1.record()
                self._gh_events.append((e0, e1))
            self.highlight = hl
            self.stats["gh_pt"] += int(pt_list.shape[0])
            self.stats["gh_ee"] += int(ee_list.shape[0])
            self.stats["gh_vg"] += int(vg_list.shape[0])
            self.stats["gh_active"] += nact
            hess = bsr_from_pairs(self.n_bodies, pairs, blocks)
            if p.solver == 3:
                bsr_ldlt(hess, self.dq, self.g, p.solver_maxiter, p.solver_tol, info=self.solve_info)
            elif p.solver in (1, 2):
                bsr_cg(hess, self.dq, self.g, p.solver_maxiter, p.solver_tol, info=self.solve_info,
                       block_precond=p.solver == 2)
            else:
                bicgstab(hess, self.g, self.dq, p.solver_tol, p.solver_maxiter, info=self.solve_info)
            alpha_cap = self.ccd(states, self.dq)
            alpha, E0 = self.line_search(alpha_cap, E0)
            self.update_q(alpha)
            it += 1
            gTg = self.dot(self.g, self.g, 0)
            self.stats["solver_iters"] += int(self.solve_info.iters.item())
            if gTg < p.tol or it >= p.max_iter:
                break
        self.update_q0qdot()
        self.update_mesh_vertex("x_view")
        self.stats["newton_iters"] += it
        self.stats["frames"] += 1
        self.E = E0
        return it

    def ccd(self, states, dq) -> float:
        """abd.py:278-303"""
        self.update_qk(1.0)
        self.update_mesh_vertex("xk")
        self.update_mesh_vertex("x")
        ij, pt_list, ee_list, vg_list = self.trajectory_intersection_set()
        self.stats["ccd_pt"] += int(pt_list.shape[0])
        self.stats["ccd_ee"] += int(ee_list.shape[0])
        t = float(toi_device(self.meshes, self.params, pt_list, ee_list, vg_list, out=self._toi).item())
        t = max(0.0, min(1.0, t))
        if t < 1.0:
            t *= 0.9
        return t

    def gh_seconds(self) -> float:
        """Sum of the gh-stage times recorded while profile_gh was set."""
        torch.cuda.synchronize()
        tot = sum(a.elapsed_time(b) for a, b in self._gh_events) * 1e-3
        self._gh_events.clear()
        return tot
