"""ctypes binding of libabd_hip.so (include/abd_hip.h).

The library is the only compute path: if it is missing the import fails
loudly (build it with `python abd-warp_amd/build.py`); there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libabd_hip.so")

ABD_OK = 0
ABD_ERR_OVERFLOW = 1
ABD_ERR_ARG = -1000
ABD_ERR_UNSORTED = -1001
ABD_ERR_NODEVICE = -1002
ABD_ERR_CAPACITY = -1003
TOI_NOALLOC_MAX_CANDIDATES = 1048576
KIND_POINT, KIND_EDGE, KIND_TRI = 0, 1, 2
CULL_EE, CULL_PT = 0, 1
SOLVER_BICGSTAB, SOLVER_CG_JACOBI, SOLVER_CG_BLOCK, SOLVER_LDLT = 0, 1, 2, 3


class AbdScene(C.Structure):
    _fields_ = [("n_bodies", C.c_int32), ("n_verts", C.c_int32), ("n_tris", C.c_int32), ("n_edges", C.c_int32),
                ("max_verts", C.c_int32), ("max_tris", C.c_int32), ("max_edges", C.c_int32)] + [
        (n, C.c_void_p) for n in ("voff", "foff", "eoff", "vert_body", "tri_body", "edge_body", "tris", "edges", "x0",
                                  "body_scene", "mass", "I0")] + [
        ("n_scenes", C.c_int32), ("scene_boff", C.c_void_p), ("scene_poff", C.c_void_p),
        ("n_scene_pairs", C.c_int64)]


class AbdSimState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("A", "p", "Ak", "pk", "A0", "p0", "Adot", "pdot", "x", "xk", "x_view", "g",
                                          "dq")]


class AbdError(RuntimeError):
    def __init__(self, fn: str, code: int):
        self.code = code
        super().__init__(f"{fn} failed: {code} ({error_string(code)})")


_lib = None


class AbdTuning(C.Structure):
    """Mirror of AbdTuning (include/abd_hip.h)."""
    _fields_ = [("cull_chunk", C.c_int32), ("cull_grid", C.c_int32), ("contact_bound_mib", C.c_int32),
                ("pair_grid", C.c_int32), ("toi_grid", C.c_int32)]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python abd-warp_amd/build.py` "
                          "(the HIP library is the only implementation of this path)")
    L = C.CDLL(LIB_PATH)
    p, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    from .const_params import AbdParamsC
    S = C.POINTER(AbdScene)    # a Structure instance passed here is sent by reference
    P = C.POINTER(AbdParamsC)
    sigs = {
        "abd_version": [],
        "abd_release": [],
        "abd_project_vertices": [p, S, p, p, p],
        "abd_update_q": [p, i32, p, p, p, d, p, p],
        "abd_update_q0qdot": [p, i32, d, p, p, p, p, p, p],
        "abd_dot": [p, p, p, i64, p],
        "abd_prim_boxes": [p, S, i32, p, d, p, p],
        "abd_prim_boxes_traj": [p, S, i32, p, p, p],
        "abd_body_boxes": [p, S, p, d, p, p],
        "abd_body_boxes_traj": [p, S, p, p, p],
        "abd_cull_bodies": [p, S, p, p, p, p, i64, p],
        "abd_cull_prims": [p, S, i32, p, i64, p, p, p, p, p, i64, p],
        "abd_cull_vg": [p, S, p, p, p, d, d, p, i64, p],
        "abd_cull_keep": [p, S, p, p, i64, p, i64, p, i64, p, i64, p, p, p, p, p, p, p, p, d, d, p, p, p, p, p,
                          C.POINTER(C.c_int64)],
        "abd_set_triplets": [p, i32, p, i64, p, p],
        "abd_bsr_structure": [p, i32, p, p, i64, p, p, p],
        "abd_contact_energy": [p, S, P, p, p, i64, p, i64, p, i64, p],
        "abd_contact_gh": [p, S, P, p, p, i64, p, i64, p, i64, i64, p, p, p, p, p, p],
        "abd_pair_gh": [p, i32, P, p, p, i64, p, p, p],
        "abd_inertia_energy": [p, S, P, p, p, p, p, p, p, p],
        "abd_inertia_gradient": [p, S, P, p, p, p, p, p, p, p],
        "abd_inertia_hessian": [p, S, P, p, p],
        "abd_bsr_solve": [p, i32, i32, p, p, p, p, p, p, d, i32, i32, p, p],
        "abd_toi": [p, S, P, p, p, p, i64, p, i64, p, i64, p],
        "abd_toi_noalloc": [p, S, P, p, p, p, i64, p, i64, p, i64, p],
        "abd_cubic_roots": [p, p, i64, d, d, p, p],
        "abd_collision_time": [p, i32, p, p, i64, p],
        "abd_scene_contact_energy": [p, S, P, p, p, i64, p, i64, p, i64, p],
        "abd_scene_inertia_energy": [p, S, P, p, p, p, p, p, p, p],
        "abd_scene_toi": [p, S, P, p, p, p, i64, p, i64, p, i64, p],
        "abd_scene_dot": [p, S, p, p, p],
        "abd_update_q_scenes": [p, S, p, p, p, p, p, p],
        "abd_bsr_solve_scenes": [p, i32, S, i32, p, p, p, p, p, p, d, i32, i32, p, p, p],
        "abd_sim_create": [S, P, C.c_void_p, C.POINTER(C.c_void_p)],
        "abd_sim_step": [p, p, C.POINTER(C.c_int64)],
        "abd_sim_scene_iters": [p, C.POINTER(C.c_int32)],
        "abd_sim_set_graphs": [p, i32],
        "abd_sim_graph_counts": [p, C.POINTER(C.c_int64)],
        "abd_sim_autotune": [p, p, C.POINTER(AbdTuning)],
        "abd_get_tuning": [C.POINTER(AbdTuning)],
        "abd_set_tuning": [C.POINTER(AbdTuning)],
        "abd_sim_hessian_solve": [p, p, i32, d, i32, p, p, C.POINTER(C.c_int32), C.POINTER(C.c_double)],
        "abd_sim_destroy": [p],
        "abd_profile_begin": [],
        "abd_profile_end": [],
        "abd_profile_select": [C.c_uint32],
        "abd_profile_query": [i32, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.abd_error_string.argtypes = [C.c_int]
    L.abd_error_string.restype = C.c_char_p
    _lib = L
    return L


def error_string(code: int) -> str:
    try:
        return load().abd_error_string(int(code)).decode()
    except Exception:  # pragma: no cover - only used for messages
        return "?"


def call(name: str, *args, ok=(ABD_OK,)) -> int:
    rc = getattr(load(), name)(*args)
    if rc not in ok:
        raise AbdError(name, rc)
    return rc


def ptr(t) -> int | None:
    """Device (or host) address of a contiguous torch tensor, None for None."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr() if t.numel() else None


def stream_handle(device=None) -> int | None:
    import torch
    s = torch.cuda.current_stream(device)
    return s.cuda_stream or None


PROF_NAMES = ["gh_classify", "gh_contact", "gh_pair_reduce", "gh_body_finalize", "energy", "ccd_toi", "cull_filter",
              "cull_count", "cull_fill", "cull_body_pairs", "cull_vg", "solve", "project", "boxes", "solve_precond",
              "ccd_toi_solve", "keep_filter"]


def profile_begin():
    call("abd_profile_begin")


def profile_end():
    call("abd_profile_end")


def profile_select(names=None):
    """Times only the kernels named (all of PROF_NAMES when None) while profiling."""
    mask = 0xFFFFFFFF if names is None else sum(1 << PROF_NAMES.index(n) for n in names)
    call("abd_profile_select", C.c_uint32(mask))


def profile_query() -> dict:
    """{kernel name: (total ms, launches, algorithmic bytes)} since profile_begin()."""
    out = {}
    for k, name in enumerate(PROF_NAMES):
        ms, n, b = C.c_double(0), C.c_int64(0), C.c_double(0)
        call("abd_profile_query", k, C.byref(ms), C.byref(n), C.byref(b))
        out[name] = (ms.value, n.value, b.value)
    return out
