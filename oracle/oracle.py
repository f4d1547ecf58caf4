"""ctypes front end of the CPU restatement (oracle/abd_oracle*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product (the abd_hip package
and libabd_hip.so) never imports it.

All arrays are host numpy in the flattened layouts of abd_hip.scenes.SceneSpec:
vertices [V,3] float64, triangles [F,3] / edges [E,2] int32 (body-local vertex
ids), states A [Nb,3,3] (row r = a_{r+1}), p [Nb,3], gradient [4Nb,3], Hessian
slots [(Nb+2Nij)*16, 3, 3] with block (ii, jj) of slot s at 16*s + ii + 4*jj.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dt", "dhat", "d2hat", "kappa", "stiffness", "gravity", "ground", "mass", "I0", "c1", "alpha_min", "tol",
        "solver_tol")] + [(n, C.c_int32) for n in ("max_iter", "solver_maxiter", "pt", "ee", "vg", "solver", "psd")]


class OrcScene(C.Structure):
    _fields_ = [("nb", C.c_int32), ("voff", C.c_void_p), ("foff", C.c_void_p), ("eoff", C.c_void_p),
                ("tris", C.c_void_p), ("edges", C.c_void_p), ("x0", C.c_void_p)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        L = _lib
        d, i, l, p = C.c_double, C.c_int, C.c_long, C.c_void_p
        L.orc_barrier.restype = d
        L.orc_barrier.argtypes = [d, d, d]
        L.orc_barrier_d1.restype = d
        L.orc_barrier_d1.argtypes = [d, d, d]
        L.orc_barrier_d2.restype = d
        L.orc_barrier_d2.argtypes = [d, d, d]
        L.orc_project.argtypes = [i, p, p, p, p, p]
        L.orc_prim_boxes.argtypes = [i, i, p, p, p, p, d, p, p]
        L.orc_prim_boxes_traj.argtypes = [i, i, p, p, p, p, p, p]
        L.orc_body_boxes.argtypes = [i, p, p, d, p, p]
        L.orc_body_boxes_traj.argtypes = [i, p, p, p, p]
        L.orc_cull_bodies.restype = l
        L.orc_cull_bodies.argtypes = [i, p, p, p, p, l]
        L.orc_cull_prims.restype = l
        L.orc_cull_prims.argtypes = [i, p, l, p, p, p, p, p, p, p, l]
        L.orc_cull_vg.restype = l
        L.orc_cull_vg.argtypes = [i, p, p, p, p, d, d, p, l]
        L.orc_contact_energy.restype = d
        L.orc_contact_energy.argtypes = [p, p, p, p, l, p, l, p, l]
        L.orc_contact_gh.argtypes = [p, p, p, p, l, p, l, p, l, l, p, p, p]
        L.orc_inertia_energy.restype = d
        L.orc_inertia_energy.argtypes = [p, i, p, p, p, p, p, p, p, p]
        L.orc_inertia_gradient.argtypes = [p, i, p, p, p, p, p, p, p, p, p]
        L.orc_inertia_hessian.argtypes = [p, i, p, p, p, p]
        L.orc_set_triplets.argtypes = [i, l, p, p, p]
        L.orc_bicgstab.restype = l
        L.orc_bicgstab.argtypes = [l, p, p, p, p, p, l, d, l, i, p]
        L.orc_bsr_cg.restype = l
        L.orc_bsr_cg.argtypes = [l, p, p, p, p, p, l, d, l, i, p]
        L.orc_bsr_cg_block.restype = l
        L.orc_bsr_cg_block.argtypes = [l, p, p, p, p, p, l, d, l, i, p]
        L.orc_bsr_ldlt.restype = l
        L.orc_bsr_ldlt.argtypes = [l, p, p, p, p, p, l]
        L.orc_cubic_roots.restype = i
        L.orc_cubic_roots.argtypes = [p, d, d, p]
        L.orc_coplanar_poly.restype = i
        L.orc_coplanar_poly.argtypes = [p, p, p]
        L.orc_collision_time_flat.restype = d
        L.orc_collision_time_flat.argtypes = [i, p, p]
        L.orc_toi.restype = d
        L.orc_toi.argtypes = [p, p, p, p, p, l, p, l, p, l]
        L.orc_pair_gh_flat.restype = i
        L.orc_pair_gh_flat.argtypes = [i, p, p, p, p, p, p, p, p]
        L.orc_full_local.argtypes = [i, p, p, d, d, i, p, p, p, p]
        L.orc_full_energy.argtypes = [i, p, p, d, d]
        L.orc_full_energy.restype = d
        L.orc_pair_geometry.argtypes = [i, p, p, p, p, p, p, p]
        L.orc_update_q.argtypes = [i, p, p, p, d, p, p]
        L.orc_update_q0qdot.argtypes = [i, d, p, p, p, p, p, p]
        L.orc_step.restype = i
        L.orc_step.argtypes = [p, p, p, p, p, i, i, i, p, p, p, p, p, p, p, p, p, p, p]
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def params(cp) -> OrcParams:
    """OrcParams from an abd_hip.const_params.ContactParams (same field order)."""
    return OrcParams(cp.dt, cp.dhat, cp.d2hat, cp.kappa, cp.stiffness, cp.gravity, cp.ground, cp.mass, cp.I0, cp.c1,
                     cp.alpha_min, cp.tol, cp.solver_tol, int(cp.max_iter), int(cp.solver_maxiter), int(bool(cp.pt)),
                     int(bool(cp.ee)), int(bool(cp.vg)), int(cp.solver), int(cp.psd))


class OracleScene:
    """Keeps the scene arrays alive for the C structure that points at them."""

    def __init__(self, spec):
        self.spec = spec
        self.voff = _c(spec.voff, np.int32)
        self.foff = _c(spec.foff, np.int32)
        self.eoff = _c(spec.eoff, np.int32)
        self.tris = _c(spec.tris, np.int32)
        self.edges = _c(spec.edges, np.int32)
        self.x0 = _c(spec.x0, np.float64)
        self.body_scene = _c(spec.body_scene, np.int32)
        self.c = OrcScene(spec.n_bodies, _ptr(self.voff), _ptr(self.foff), _ptr(self.eoff), _ptr(self.tris),
                          _ptr(self.edges), _ptr(self.x0))
        self.nb = spec.n_bodies


# ----------------------------------------------------------- stage wrappers
def barrier(s, d2hat, kappa):
    return lib().orc_barrier(float(s), float(d2hat), float(kappa))


def barrier_d1(s, d2hat, kappa):
    return lib().orc_barrier_d1(float(s), float(d2hat), float(kappa))


def barrier_d2(s, d2hat, kappa):
    return lib().orc_barrier_d2(float(s), float(d2hat), float(kappa))


def project(sc: OracleScene, A, p):
    A = _c(A, np.float64)
    p = _c(p, np.float64)
    x = np.zeros((sc.spec.n_verts, 3))
    lib().orc_project(sc.nb, _ptr(sc.voff), _ptr(sc.x0), _ptr(A), _ptr(p), _ptr(x))
    return x


KIND_POINT, KIND_EDGE, KIND_TRI = 0, 1, 2


def prim_boxes(sc: OracleScene, kind: int, x, dil: float):
    x = _c(x, np.float64)
    n = {0: sc.spec.n_verts, 1: sc.spec.n_edges, 2: sc.spec.n_tris}[kind]
    offs = {0: sc.voff, 1: sc.eoff, 2: sc.foff}[kind]
    topo = {0: None, 1: sc.edges, 2: sc.tris}[kind]
    lo = np.zeros((n, 3), np.float32)
    hi = np.zeros((n, 3), np.float32)
    lib().orc_prim_boxes(kind, sc.nb, _ptr(sc.voff), _ptr(offs), _ptr(topo), _ptr(x), float(dil), _ptr(lo), _ptr(hi))
    return lo, hi


def prim_boxes_traj(sc: OracleScene, kind: int, x_end, lo, hi):
    x_end = _c(x_end, np.float64)
    lo = np.array(lo, np.float32, copy=True)
    hi = np.array(hi, np.float32, copy=True)
    offs = {0: sc.voff, 1: sc.eoff, 2: sc.foff}[kind]
    topo = {0: None, 1: sc.edges, 2: sc.tris}[kind]
    lib().orc_prim_boxes_traj(kind, sc.nb, _ptr(sc.voff), _ptr(offs), _ptr(topo), _ptr(x_end), _ptr(lo), _ptr(hi))
    return lo, hi


def body_boxes(sc: OracleScene, x, dil: float):
    x = _c(x, np.float64)
    lo = np.zeros((sc.nb, 3), np.float32)
    hi = np.zeros((sc.nb, 3), np.float32)
    lib().orc_body_boxes(sc.nb, _ptr(sc.voff), _ptr(x), float(dil), _ptr(lo), _ptr(hi))
    return lo, hi


def body_boxes_traj(sc: OracleScene, x_end, lo, hi):
    x_end = _c(x_end, np.float64)
    lo = np.array(lo, np.float32, copy=True)
    hi = np.array(hi, np.float32, copy=True)
    lib().orc_body_boxes_traj(sc.nb, _ptr(sc.voff), _ptr(x_end), _ptr(lo), _ptr(hi))
    return lo, hi


def cull_bodies(sc: OracleScene, lo, hi, use_scenes: bool = True):
    lo = _c(lo, np.float32)
    hi = _c(hi, np.float32)
    bs = sc.body_scene if use_scenes else None
    n = lib().orc_cull_bodies(sc.nb, _ptr(bs), _ptr(lo), _ptr(hi), None, 0)
    ij = np.zeros((max(n, 1), 2), np.int32)
    lib().orc_cull_bodies(sc.nb, _ptr(bs), _ptr(lo), _ptr(hi), _ptr(ij), n)
    return ij[:n]


def cull_prims(pt: bool, ij, off1, lo1, hi1, off2, lo2, hi2):
    ij = _c(ij, np.int32)
    args = [_c(a, t) for a, t in ((off1, np.int32), (lo1, np.float32), (hi1, np.float32), (off2, np.int32),
                                   (lo2, np.float32), (hi2, np.float32))]
    ptrs = [_ptr(a) for a in args]
    n = lib().orc_cull_prims(int(pt), _ptr(ij), ij.shape[0], *ptrs, None, 0)
    out = np.zeros((max(n, 1), 5), np.int32)
    lib().orc_cull_prims(int(pt), _ptr(ij), ij.shape[0], *ptrs, _ptr(out), n)
    return out[:n]


def cull_vg(sc: OracleScene, blo, plo, phi, dhat, ground):
    blo, plo, phi = (_c(a, np.float32) for a in (blo, plo, phi))
    n = lib().orc_cull_vg(sc.nb, _ptr(sc.voff), _ptr(blo), _ptr(plo), _ptr(phi), float(dhat), float(ground), None, 0)
    out = np.zeros((max(n, 1), 2), np.int32)
    lib().orc_cull_vg(sc.nb, _ptr(sc.voff), _ptr(blo), _ptr(plo), _ptr(phi), float(dhat), float(ground), _ptr(out), n)
    return out[:n]


def _lists(pt, ee, vg):
    pt = _c(np.zeros((0, 5)) if pt is None else pt, np.int32)
    ee = _c(np.zeros((0, 5)) if ee is None else ee, np.int32)
    vg = _c(np.zeros((0, 2)) if vg is None else vg, np.int32)
    return pt, ee, vg


def contact_energy(sc: OracleScene, prm: OrcParams, xk, pt, ee, vg):
    xk = _c(xk, np.float64)
    pt, ee, vg = _lists(pt, ee, vg)
    return lib().orc_contact_energy(C.byref(sc.c), C.byref(prm), _ptr(xk), _ptr(pt), pt.shape[0], _ptr(ee),
                                    ee.shape[0], _ptr(vg), vg.shape[0])


def contact_gh(sc: OracleScene, prm: OrcParams, x, pt, ee, vg, nij, g=None, blocks=None):
    """Returns (g [4Nb,3], blocks [(Nb+2Nij)*16,3,3], highlight uint8 [Nb])."""
    x = _c(x, np.float64)
    pt, ee, vg = _lists(pt, ee, vg)
    nb = sc.nb
    g = np.zeros((4 * nb, 3)) if g is None else np.array(g, np.float64, copy=True)
    blocks = np.zeros(((nb + 2 * nij) * 16, 3, 3)) if blocks is None else np.array(blocks, np.float64, copy=True)
    hl = np.zeros(nb, np.uint8)
    lib().orc_contact_gh(C.byref(sc.c), C.byref(prm), _ptr(x), _ptr(pt), pt.shape[0], _ptr(ee), ee.shape[0],
                         _ptr(vg), vg.shape[0], int(nij), _ptr(g), _ptr(blocks), _ptr(hl))
    return g, blocks, hl


def pair_gh(kind: int, prm: OrcParams, x4, r4):
    x4 = _c(x4, np.float64)
    r4 = _c(r4, np.float64)
    g24 = np.zeros(24)
    Hi, Hj, Hij = np.zeros((12, 12)), np.zeros((12, 12)), np.zeros((12, 12))
    s = np.zeros(1)
    act = lib().orc_pair_gh_flat(kind, C.byref(prm), _ptr(x4), _ptr(r4), _ptr(g24), _ptr(Hi), _ptr(Hj), _ptr(Hij),
                                 _ptr(s))
    return bool(act), float(s[0]), g24, Hi, Hj, Hij


def full_local(kind: int, x4, r4, d2hat: float, kappa: float, project: bool = True):
    """Full-IPC mode, one candidate in its 12 point coordinates: (active, E, g12, H12, feature)
    with feature = (type, i, j, k): type 0 point-plane / line-line, 1 point x[i] - edge x[j] x[k],
    2 point x[i] - point x[j]."""
    x4 = _c(x4, np.float64)
    r4 = _c(r4, np.float64)
    E = np.zeros(1)
    g, H, f = np.zeros(12), np.zeros((12, 12)), np.zeros(4, np.int32)
    act = lib().orc_full_local(kind, _ptr(x4), _ptr(r4), d2hat, kappa, 1 if project else 0, _ptr(E), _ptr(g),
                               _ptr(H), _ptr(f))
    return bool(act), float(E[0]), g, H, tuple(int(v) for v in f)


def full_energy(kind: int, x4, r4, d2hat: float, kappa: float) -> float:
    return float(lib().orc_full_energy(kind, _ptr(_c(x4, np.float64)), _ptr(_c(r4, np.float64)), d2hat, kappa))


def pair_geometry(kind: int, x4):
    x4 = _c(x4, np.float64)
    c9, bgl = np.zeros(9), np.zeros(3)
    valid = C.c_int(0)
    Q, lam, D = np.zeros(45), np.zeros(5), np.zeros(12)
    lib().orc_pair_geometry(kind, _ptr(x4), _ptr(c9), _ptr(bgl), C.byref(valid), _ptr(Q), _ptr(lam), _ptr(D))
    return dict(e0p=c9[:3], e1p=c9[3:6], e2p=c9[6:], beta=bgl[0], gamma=bgl[1], l=bgl[2], valid=bool(valid.value),
                Q=Q.reshape(5, 9), lam=lam, D=D.reshape(3, 4))


def _states(st):
    return [_c(st[k], np.float64) for k in ("A", "p", "A0", "p0", "Adot", "pdot")]


def inertia_energy(prm: OrcParams, Ak, pk, A0, p0, Adot, pdot, mass=None, I0=None):
    arrs = [_c(a, np.float64) for a in (Ak, pk, A0, p0, Adot, pdot)]
    m = None if mass is None else _c(mass, np.float64)
    i0 = None if I0 is None else _c(I0, np.float64)
    return lib().orc_inertia_energy(C.byref(prm), arrs[0].shape[0], *[_ptr(a) for a in arrs], _ptr(m), _ptr(i0))


def inertia_gradient(prm: OrcParams, A, p, A0, p0, Adot, pdot, mass=None, I0=None):
    arrs = [_c(a, np.float64) for a in (A, p, A0, p0, Adot, pdot)]
    m = None if mass is None else _c(mass, np.float64)
    i0 = None if I0 is None else _c(I0, np.float64)
    g = np.zeros((4 * arrs[0].shape[0], 3))
    lib().orc_inertia_gradient(C.byref(prm), arrs[0].shape[0], *[_ptr(a) for a in arrs], _ptr(m), _ptr(i0), _ptr(g))
    return g


def inertia_hessian(prm: OrcParams, A, blocks, mass=None, I0=None):
    A = _c(A, np.float64)
    blocks = np.array(blocks, np.float64, copy=True)
    m = None if mass is None else _c(mass, np.float64)
    i0 = None if I0 is None else _c(I0, np.float64)
    lib().orc_inertia_hessian(C.byref(prm), A.shape[0], _ptr(A), _ptr(m), _ptr(i0), _ptr(blocks))
    return blocks


def set_triplets(nb: int, ij):
    ij = _c(ij, np.int32)
    nij = ij.shape[0]
    rows = np.zeros((nb + 2 * nij) * 16, np.int32)
    cols = np.zeros_like(rows)
    lib().orc_set_triplets(nb, nij, _ptr(ij), _ptr(rows), _ptr(cols))
    return rows, cols


def bicgstab(nb: int, ij, blocks, b, x, tol, maxiter=0, check_every=10):
    rows, cols = set_triplets(nb, ij)
    blocks = _c(blocks, np.float64)
    b = _c(b, np.float64)
    x = np.array(x, np.float64, copy=True)
    res = np.zeros(1)
    it = lib().orc_bicgstab(blocks.shape[0] // 16, _ptr(rows), _ptr(cols), _ptr(blocks), _ptr(b), _ptr(x),
                            b.size, float(tol), int(maxiter), int(check_every), _ptr(res))
    return x, int(it), float(res[0])


def bsr_cg(nb: int, ij, blocks, b, x, tol, maxiter=0, check_every=10):
    rows, cols = set_triplets(nb, ij)
    blocks = _c(blocks, np.float64)
    b = _c(b, np.float64)
    x = np.array(x, np.float64, copy=True)
    res = np.zeros(1)
    it = lib().orc_bsr_cg(blocks.shape[0] // 16, _ptr(rows), _ptr(cols), _ptr(blocks), _ptr(b), _ptr(x), b.size,
                          float(tol), int(maxiter), int(check_every), _ptr(res))
    return x, int(it), float(res[0])


def bsr_cg_block(nb: int, ij, blocks, b, x, tol, maxiter=0, check_every=10):
    """CG with the per-body 12x12 diagonal block as preconditioner (beyond the reference)."""
    rows, cols = set_triplets(nb, ij)
    blocks = _c(blocks, np.float64)
    b = _c(b, np.float64)
    x = np.array(x, np.float64, copy=True)
    res = np.zeros(1)
    it = lib().orc_bsr_cg_block(blocks.shape[0] // 16, _ptr(rows), _ptr(cols), _ptr(blocks), _ptr(b), _ptr(x),
                                b.size, float(tol), int(maxiter), int(check_every), _ptr(res))
    return x, int(it), float(res[0])


def bsr_ldlt(nb: int, ij, blocks, b, x):
    """Direct block LDL^T solve (beyond the reference); returns (x, status): status 1, or -2
    with x unchanged when a pivot is not positive or the system has more than 64 bodies."""
    rows, cols = set_triplets(nb, ij)
    blocks = _c(blocks, np.float64)
    b = _c(b, np.float64)
    x = np.array(x, np.float64, copy=True)
    st = lib().orc_bsr_ldlt(blocks.shape[0] // 16, _ptr(rows), _ptr(cols), _ptr(blocks), _ptr(b), _ptr(x), b.size)
    return x, int(st)


def cubic_roots(coef, x0=0.0, x1=1.0):
    coef = _c(coef, np.float64)
    roots = np.zeros(3)
    n = lib().orc_cubic_roots(_ptr(coef), float(x0), float(x1), _ptr(roots))
    return n, roots


def coplanar_poly(p_t0, p_t1):
    a = _c(p_t0, np.float64)
    b = _c(p_t1, np.float64)
    coef = np.zeros(4)
    lib().orc_coplanar_poly(_ptr(a), _ptr(b), _ptr(coef))
    return coef


def collision_time(kind: int, p_t0, p_t1):
    return lib().orc_collision_time_flat(kind, _ptr(_c(p_t0, np.float64)), _ptr(_c(p_t1, np.float64)))


def toi(sc: OracleScene, prm: OrcParams, x, xk, pt, ee, vg):
    x = _c(x, np.float64)
    xk = _c(xk, np.float64)
    pt, ee, vg = _lists(pt, ee, vg)
    return lib().orc_toi(C.byref(sc.c), C.byref(prm), _ptr(x), _ptr(xk), _ptr(pt), pt.shape[0], _ptr(ee),
                         ee.shape[0], _ptr(vg), vg.shape[0])


class OracleSim:
    """The oracle's frame driver (abd.py:212-276) over a SceneSpec."""

    def __init__(self, spec, cp):
        self.sc = OracleScene(spec)
        self.prm = params(cp)
        nb = spec.n_bodies
        self.A0 = _c(spec.A, np.float64).copy()
        self.p0 = _c(spec.p, np.float64).copy()
        self.Adot = _c(spec.Adot, np.float64).copy()
        self.pdot = _c(spec.pdot, np.float64).copy()
        self.A = self.A0.copy()
        self.p = self.p0.copy()
        self.Ak = self.A0.copy()
        self.pk = self.p0.copy()
        self.dq = np.zeros((4 * nb, 3))
        self.mass = _c(spec.mass, np.float64)
        self.I0 = _c(spec.I0, np.float64)
        self.stats = np.zeros(8, np.int64)
        self.E = 0.0

    def step(self) -> int:
        s = self.sc
        E = C.c_double(0.0)
        arrs = [self.A, self.p, self.Ak, self.pk, self.A0, self.p0, self.Adot, self.pdot, self.dq]
        it = lib().orc_step(C.byref(s.c), C.byref(self.prm), _ptr(s.body_scene), _ptr(self.mass), _ptr(self.I0),
                            s.spec.n_verts, s.spec.n_tris, s.spec.n_edges, *[_ptr(a) for a in arrs],
                            _ptr(self.stats), C.byref(E))
        self.E = E.value
        return it
