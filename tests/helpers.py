"""Shared test utilities: seeded contact geometries and oracle/GPU glue."""
from __future__ import annotations

import numpy as np

from abd_hip.const_params import ContactParams
from oracle import oracle as O


def rel_err(a, b) -> float:
    """max |a - b| / max(max |b|, tiny): the norm-wise relative error used for
    the 1e-9 Hessian / gradient / energy bound."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def near_contact_pt(rng, dist):
    """A point at signed distance `dist` above a random triangle, projecting inside it."""
    while True:
        t = rng.uniform(-1, 1, size=(3, 3))
        e0, e1 = t[1] - t[0], t[2] - t[0]
        n = np.cross(e0, e1)
        if np.linalg.norm(n) < 0.3:
            continue
        n /= np.linalg.norm(n)
        b = rng.uniform(0.05, 0.9)
        g = rng.uniform(0.05, 0.95 - b)
        p = t[0] + b * e0 + g * e1 + dist * n
        return np.stack([p, t[0], t[1], t[2]])


def near_contact_ee(rng, dist):
    """Two skew edges whose closest points are interior, at distance `dist`."""
    while True:
        a0 = rng.uniform(-1, 1, 3)
        d0 = rng.uniform(-1, 1, 3)
        d1 = rng.uniform(-1, 1, 3)
        n = np.cross(d0, d1)
        if np.linalg.norm(n) < 0.3:
            continue
        n /= np.linalg.norm(n)
        s, t = rng.uniform(0.15, 0.85, 2)
        c0 = a0 + s * d0
        c1 = c0 + dist * n
        b0 = c1 - t * d1
        return np.stack([a0, a0 + d0, b0, b0 + d1])


def rest_of(x4, rng, scale=0.3):
    """Rest positions: a random affine pre-image-like perturbation of x4."""
    return x4 + scale * rng.uniform(-1, 1, size=x4.shape)


def params(**kw) -> ContactParams:
    return ContactParams(**kw)


def orc_params(cp: ContactParams):
    return O.params(cp)


def rot_y(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])


def rot_axis(axis, a):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K


def contact_scene(gap=4e-4, dhat=1e-3, seed=0):
    """Four bodies with active PT, EE and VG contacts at distance ~gap < dhat:
    cube 0 on the ground, cube 1 (rotated about y) on cube 0, an icosphere-2 on the
    ground, a tilted icosphere-2 against cube 0's +x face.  A holds rotations as
    rows, so vertices map to x = A^T x0 + p."""
    from abd_hip import scenes
    rng = np.random.default_rng(seed)
    cube = scenes.cube_mesh()
    ico = scenes.icosphere_mesh(2, 0.5)
    ground = -0.5 - gap
    R1 = rot_y(0.3 + 0.01 * rng.uniform())
    R3 = rot_axis([0.3, 1.0, 0.2], 0.7)
    A = np.stack([np.eye(3), R1.T, np.eye(3), R3.T])  # x = A^T x0 + p = R x0 + p
    ico3 = ico[0] @ R3.T
    p = np.array([
        [0.0, 0.0, 0.0],
        [0.05, 1.0 + gap, -0.03],
        [2.0, ground + gap - ico[0][:, 1].min(), 0.0],
        [0.5 + gap - ico3[:, 0].min(), 0.0, 0.02],
    ])
    z = np.zeros((4, 3))
    return scenes.build_scene([cube, cube, ico, ico], A, p, np.zeros((4, 3, 3)), z, ground=ground, name="contact")


def oracle_lists(sc, x, dhat, ground, traj_from=None):
    """The oracle's collision_set (abd.py:187-207) at positions x: boxes, ij, ee, pt, vg."""
    dh = 0.5 * dhat
    blo, bhi = O.body_boxes(sc, x, dh)
    tlo, thi = O.prim_boxes(sc, O.KIND_TRI, x, 0.0)
    elo, ehi = O.prim_boxes(sc, O.KIND_EDGE, x, dh)
    plo, phi = O.prim_boxes(sc, O.KIND_POINT, x, dhat)
    ij = O.cull_bodies(sc, blo, bhi)
    ee = O.cull_prims(False, ij, sc.eoff, elo, ehi, sc.eoff, elo, ehi)
    pt = O.cull_prims(True, ij, sc.voff, plo, phi, sc.foff, tlo, thi)
    vg = O.cull_vg(sc, blo, plo, phi, dhat, ground)
    return dict(blo=blo, bhi=bhi, tlo=tlo, thi=thi, elo=elo, ehi=ehi, plo=plo, phi=phi, ij=ij, ee=ee, pt=pt, vg=vg)


def _unit(v):
    return v / np.linalg.norm(v)


def full_ipc_config(rng, kind: int, feature: str, dist: float):
    """A PT (kind 0) or EE (kind 1) contact whose closest feature pair is `feature`
    ('face' / 'line', 'edge' / 'pe', 'vertex' / 'pp', 'parallel'), at distance `dist`."""
    while True:
        t = rng.uniform(-1, 1, size=(3, 3))
        if np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])) > 0.5:
            break
    if kind == 0:
        n = _unit(np.cross(t[1] - t[0], t[2] - t[0]))
        if feature == "face":
            return near_contact_pt(rng, dist)
        if feature == "edge":
            k = rng.integers(3)
            a, b, c = t[k], t[(k + 1) % 3], t[(k + 2) % 3]
            e = _unit(b - a)
            out = _unit(np.cross(e, n))
            if np.dot(out, c - a) > 0:
                out = -out
            ang = rng.uniform(-1.0, 1.0)
            d = _unit(np.cos(ang) * out + np.sin(ang) * n)
            p = a + rng.uniform(0.2, 0.8) * (b - a) + dist * d
            return np.stack([p, t[0], t[1], t[2]])
        k = rng.integers(3)  # vertex: beyond the vertex, away from the triangle
        a = t[k]
        away = _unit(2 * a - t[(k + 1) % 3] - t[(k + 2) % 3])
        d = _unit(away + rng.uniform(-0.3, 0.3) * n)
        return np.stack([a + dist * d, t[0], t[1], t[2]])
    if feature == "line":
        return near_contact_ee(rng, dist)
    b0 = rng.uniform(-1, 1, 3)
    db = _unit(rng.uniform(-1, 1, 3))
    b1 = b0 + rng.uniform(0.8, 1.5) * db
    perp = _unit(np.cross(db, rng.uniform(-1, 1, 3)))
    if feature == "pe":  # endpoint a0 near the interior of edge b, edge a leaving it
        a0 = b0 + rng.uniform(0.3, 0.7) * (b1 - b0) + dist * perp
        da = _unit(perp + rng.uniform(-0.4, 0.4, 3))
        return np.stack([a0, a0 + rng.uniform(0.6, 1.2) * da, b0, b1])
    if feature == "pp":  # endpoints a0 and b0 close, both edges leaving
        a0 = b0 + dist * _unit(-db + rng.uniform(-0.3, 0.3, 3))
        da = _unit(-db + rng.uniform(-0.5, 0.5, 3))
        return np.stack([a0, a0 + rng.uniform(0.6, 1.2) * da, b0, b1])
    # nearly parallel, overlapping, closest points inside both edges
    da = _unit(db + 0.01 * perp + 0.01 * _unit(np.cross(db, perp)))
    off = _unit(np.cross(da, db)) if np.linalg.norm(np.cross(da, db)) > 0 else perp
    a0 = b0 + 0.25 * (b1 - b0) + dist * off - 0.1 * da
    return np.stack([a0, a0 + 0.7 * np.linalg.norm(b1 - b0) * da, b0, b1])


# ------------------------------------------------------------------ driver pins
# Scenes whose frame has a closed-form outcome (abd.py:212-303): a free-falling
# cube, a cube falling onto the ground, a cube falling onto another.  Each case
# runs one frame of any driver (oracle or GPU) and is checked against
# arithmetic written out here, not against another implementation.

def one_cube(p, pdot, ground=-100.0):
    from abd_hip import scenes
    V, F = scenes.cube_mesh()
    return scenes.build_scene([(V, F)], np.eye(3)[None], np.asarray([p], np.float64), np.zeros((1, 3, 3)),
                              np.asarray([pdot], np.float64), ground=ground, name="one-cube")


def two_cubes(pA, pdotA, pB, ground=-100.0):
    from abd_hip import scenes
    V, F = scenes.cube_mesh()
    return scenes.build_scene([(V, F), (V, F)], np.stack([np.eye(3)] * 2), np.asarray([pB, pA], np.float64),
                              np.zeros((2, 3, 3)), np.asarray([[0.0, 0.0, 0.0], pdotA], np.float64), ground=ground,
                              name="two-cubes")


def inertial_target(spec, cp):
    """p~ = p0 + dt pdot + dt^2 (0, gravity, 0) (orthogonal_energy.py:25-52)."""
    return spec.p + cp.dt * spec.pdot + cp.dt * cp.dt * np.array([0.0, cp.gravity, 0.0])


def _planted(spec, cp, scale):
    """Warm start dq = scale * (p0 - p~) on the first body's translation (zero on A)."""
    dq = np.zeros((4 * spec.n_bodies, 3))
    dq[0] = scale * (spec.p[0] - inertial_target(spec, cp)[0])
    return dq


def free_energy(spec, cp, p):
    """Inertial energy of a contact-free translated cube at A = I (orthogonal_energy.py:25-52)."""
    d = p - inertial_target(spec, cp)[0]
    return 0.5 * cp.mass * float((d * d).sum())


def driver_pin_cases():
    """{name: (spec, params, planted warm start or None, expected p after one frame,
    expected Newton iterations, tolerance on p)}.  The line-search cases plant the
    search direction through the warm start: with solver_tol so large that the Jacobi
    CG's first residual test passes (bsr_utils.py:195-205), the solve returns it."""
    from abd_hip.const_params import ContactParams
    cases = {}
    hi = [0.0, 5.0, 0.0]
    v = [0.3, -0.7, 0.2]
    # free fall: the inertial energy is quadratic in p with its minimum at p~ and A = I is a
    # stationary point of the orthogonality energy; iteration 1 lands on p~ (g.g = m^2 |p0 - p~|^2
    # > tol), iteration 2 sees g = 0 and stops
    for solver in (0, 1, 2, 3):
        spec = one_cube(hi, v)
        cp = ContactParams(ground=spec.ground, solver=solver)
        cases[f"free_fall_solver{solver}"] = (spec, cp, None, inertial_target(spec, cp), 2, 1e-12)
    # the stop test reads g before the update: a tolerance above the first g.g still takes
    # that iteration's full step (at least one iteration per frame)
    spec = one_cube(hi, v)
    cp = ContactParams(ground=spec.ground, solver=3, tol=1e30)
    cases["stop_after_first_update"] = (spec, cp, None, inertial_target(spec, cp), 1, 1e-12)
    spec = one_cube(hi, v)
    cp = ContactParams(ground=spec.ground, solver=3, max_iter=1)
    cases["max_iter_1"] = (spec, cp, None, inertial_target(spec, cp), 1, 1e-12)
    # D8: E1 < E0 + c1 alpha q.g with the step -alpha dq and q.g > 0 accepts an increase below
    # c1 alpha q.g at alpha = 1 (E(1) = (1 + 1e-4)^2 E0 here)
    spec = one_cube(hi, v)
    cp = ContactParams(ground=spec.ground, solver=1, solver_tol=1e30, max_iter=1)
    dq = _planted(spec, cp, 2.0 + 1e-4)
    cases["armijo_as_written_accepts"] = (spec, cp, dq, spec.p - dq[:1], 1, 1e-12)
    # E(1) = 1.0201 E0, above the bound: one halving
    dq = _planted(spec, cp, 2.0 + 1e-2)
    cases["armijo_halves"] = (spec, cp, dq, spec.p - 0.5 * dq[:1], 1, 1e-12)
    # an ascent direction is never accepted: the halving stops at the first step length
    # below alpha_min (2^-10, evaluated before the test); the planted direction survives
    # every solve, so each of max_iter = 4 iterations takes that step
    cp4 = ContactParams(ground=spec.ground, solver=1, solver_tol=1e30, max_iter=4)
    dq = _planted(spec, cp4, -1.0)
    cases["alpha_min_and_max_iter"] = (spec, cp4, dq, spec.p - 4 * 2.0 ** -10 * dq[:1], 4, 1e-12)
    # CCD cap against the ground: bottom face 0.05 above it at 10 m/s down; the vertex-ground
    # TOI is d0 / (d0 - d1) (ccd.py:189-195), the step 0.9 of it (abd.py:300-302), leaving a
    # gap of 0.1 d0 > dhat (no barrier), which the line search accepts
    spec = one_cube([0.0, 0.55, 0.0], [0.0, -10.0, 0.0], ground=0.0)
    cp = ContactParams(ground=0.0, solver=3, max_iter=1)
    pt = inertial_target(spec, cp)
    d0 = 0.05
    d1 = d0 + (pt[0, 1] - spec.p[0, 1])
    a = 0.9 * d0 / (d0 - d1)
    cases["ccd_cap_ground"] = (spec, cp, None, spec.p + a * (pt - spec.p), 1, 1e-12)
    # CCD cap between two cubes: A shifted in x and z so that vertices of each face lie
    # strictly inside the other's face, falling at 5 m/s onto B; the PT / EE TOIs are the time
    # the relative vertical motion closes the gap (cubic roots bracketed to 1e-10,
    # cubic_roots.py:4)
    gap = 0.04
    spec = two_cubes([0.1, 1.0 + gap, 0.13], [0.0, -5.0, 0.0], [0.0, 0.0, 0.0])
    cp = ContactParams(ground=spec.ground, solver=3, max_iter=1)
    pt = inertial_target(spec, cp)
    rel = (pt[1, 1] - spec.p[1, 1]) - (pt[0, 1] - spec.p[0, 1])
    a = 0.9 * gap / (gap - (gap + rel))
    cases["ccd_cap_two_cubes"] = (spec, cp, None, spec.p + a * (pt - spec.p), 1, 1e-9)
    return cases
