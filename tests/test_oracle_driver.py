"""The oracle's frame driver (abd.py:212-276).

Closed-form pins of its control logic, each chosen so that a plausible mistake
fails it: the implicit-Euler free fall (the inertial target p~ = p0 + dt pdot +
dt^2 g, the update sign, the pre-update g.g stopping test, at least one
iteration per frame, the max_iter cutoff); the CCD step cap 0.9 d0/(d0 - d1)
against the ground and between two cubes (abd.py:300-302); the line search's
acceptance test E1 < E0 + c1 alpha q.g as written (D8: an energy increase
smaller than c1 alpha q.g is accepted), the halving and the alpha_min cutoff
(abd.py:146-160).  The search direction of the line-search cases is planted
through the warm start: with solver_tol so large that the Jacobi CG's first
residual test passes (bsr_utils.py:195-205), the solve returns the warm start
unchanged.  Then physical sanity on the two-cube scene and determinism.
"""
import numpy as np
import pytest

from abd_hip import scenes
from abd_hip.const_params import ContactParams
from helpers import inertial_target, one_cube, two_cubes
from oracle import oracle as O

HI = [0.0, 5.0, 0.0]


def _frame(spec, cp, dq=None):
    sim = O.OracleSim(spec, cp)
    if dq is not None:
        sim.dq[...] = dq
    it = sim.step()
    return sim, it


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_free_fall_is_one_newton_step(solver):
    """Contact-free cube: the inertial energy is quadratic in p with its minimum at
    p~, A = I is a stationary point of the orthogonality energy, so the first
    Newton step lands on p~ and the second iteration sees g = 0 and stops."""
    spec = one_cube(HI, [0.3, -0.7, 0.2])
    cp = ContactParams(ground=spec.ground, solver=solver)
    sim, it = _frame(spec, cp)
    pt = inertial_target(spec, cp)
    assert it == 2          # iteration 1 has g.g = m^2 |p0 - p~|^2 > tol; iteration 2 has g.g ~ 0
    np.testing.assert_allclose(sim.p, pt, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(sim.A, np.eye(3)[None])
    np.testing.assert_allclose(sim.pdot, (pt - spec.p) / cp.dt, rtol=0, atol=1e-10)
    np.testing.assert_array_equal(sim.p0, sim.p)


def test_stop_test_reads_the_gradient_before_the_update():
    """With a tolerance above g.g of the first iteration the frame still takes
    that iteration's full step, then stops (at least one iteration per frame,
    test after the update of q)."""
    spec = one_cube(HI, [0.3, -0.7, 0.2])
    cp = ContactParams(ground=spec.ground, solver=3, tol=1e30)
    sim, it = _frame(spec, cp)
    assert it == 1
    np.testing.assert_allclose(sim.p, inertial_target(spec, cp), rtol=0, atol=1e-12)


def test_max_iter_cuts_the_loop():
    spec = one_cube(HI, [0.3, -0.7, 0.2])
    cp = ContactParams(ground=spec.ground, solver=3, max_iter=1)
    sim, it = _frame(spec, cp)
    assert it == 1
    np.testing.assert_allclose(sim.p, inertial_target(spec, cp), rtol=0, atol=1e-12)


def _planted(spec, cp, scale):
    """Warm start dq = scale * (p0 - p~) on the cube's translation (zero on A)."""
    dq = np.zeros((4, 3))
    dq[0] = scale * (spec.p[0] - inertial_target(spec, cp)[0])
    return dq


def _energy(spec, cp, p):
    # inertial energy of a contact-free translated cube at A = I (orthogonal_energy.py:25-52)
    d = p - inertial_target(spec, cp)
    return 0.5 * cp.mass * float((d * d).sum())


def test_line_search_accepts_a_small_increase_as_written():
    """D8: the test is E1 < E0 + c1 alpha q.g with the step -alpha dq and q.g > 0,
    so an energy increase below c1 alpha q.g is accepted at alpha = 1 (a test
    E1 < E0 - c1 alpha q.g, or E1 < E0, would halve)."""
    spec = one_cube(HI, [0.3, -0.7, 0.2])
    cp = ContactParams(ground=spec.ground, solver=1, solver_tol=1e30, max_iter=1)
    dq = _planted(spec, cp, 2.0 + 1e-4)
    sim, it = _frame(spec, cp, dq)
    p1 = spec.p[0] - dq[0]
    E0, E1 = _energy(spec, cp, spec.p[0]), _energy(spec, cp, p1)
    g = cp.mass * (spec.p[0] - inertial_target(spec, cp)[0])
    qTg = float(dq[0] @ g)
    assert E0 < E1 < E0 + cp.c1 * qTg           # the case is the D8 branch
    assert it == 1
    np.testing.assert_allclose(sim.p[0], p1, rtol=0, atol=1e-12)
    assert sim.E == pytest.approx(E1, rel=1e-12)


def test_line_search_halves_past_the_armijo_bound():
    spec = one_cube(HI, [0.3, -0.7, 0.2])
    cp = ContactParams(ground=spec.ground, solver=1, solver_tol=1e30, max_iter=1)
    dq = _planted(spec, cp, 2.0 + 1e-2)         # E(1) = 1.0201 E0: above E0 + c1 q.g
    sim, _ = _frame(spec, cp, dq)
    np.testing.assert_allclose(sim.p[0], spec.p[0] - 0.5 * dq[0], rtol=0, atol=1e-12)


def test_line_search_stops_below_alpha_min_and_max_iter_bounds_the_frame():
    """An ascent direction is never accepted: the halving runs until alpha <
    alpha_min (1e-3), so the step taken is 2^-10 (the first step length below
    alpha_min, evaluated before the test); with the planted direction kept by
    every solve, each of the max_iter iterations takes that step."""
    spec = one_cube(HI, [0.3, -0.7, 0.2])
    cp = ContactParams(ground=spec.ground, solver=1, solver_tol=1e30, max_iter=4)
    dq = _planted(spec, cp, -1.0)
    sim, it = _frame(spec, cp, dq)
    assert it == 4
    np.testing.assert_allclose(sim.p[0], spec.p[0] - 4 * 2.0 ** -10 * dq[0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(sim.dq, dq, rtol=0, atol=0)


def test_ground_ccd_caps_the_first_step():
    """A cube 0.05 above the ground at 10 m/s downward: the full Newton step
    would pass through the ground; the vertex-ground TOI of the bottom face is
    t = d0 / (d0 - d1) (ccd.py:189-195) and the step is capped at 0.9 t, which
    leaves a gap of 0.1 d0 (> dhat: no barrier), so the line search accepts it."""
    ground = 0.0
    spec = one_cube([0.0, ground + 0.5 + 0.05, 0.0], [0.0, -10.0, 0.0], ground=ground)
    cp = ContactParams(ground=ground, solver=3, max_iter=1)
    sim, it = _frame(spec, cp)
    pt = inertial_target(spec, cp)[0]
    d0 = 0.05
    d1 = d0 + (pt[1] - spec.p[0, 1])
    alpha = 0.9 * d0 / (d0 - d1)
    assert it == 1
    np.testing.assert_allclose(sim.p[0], spec.p[0] + alpha * (pt - spec.p[0]), rtol=0, atol=1e-12)


def test_two_cube_ccd_caps_the_first_step():
    """Cube A (shifted in x and z so that vertices of each face lie strictly
    inside the other's face) falls onto cube B at 5 m/s: the point-triangle and
    edge-edge TOIs of the touching faces are the time the relative vertical
    motion closes the gap d0, the cubic roots of ccd.py:134-186 (bracketed to
    1e-10, cubic_roots.py:4)."""
    gap = 0.04
    pA = [0.1, 1.0 + gap, 0.13]
    spec = two_cubes(pA, [0.0, -5.0, 0.0], [0.0, 0.0, 0.0])
    cp = ContactParams(ground=spec.ground, solver=3, max_iter=1)
    sim, it = _frame(spec, cp)
    pt = inertial_target(spec, cp)
    rel = (pt[1, 1] - spec.p[1, 1]) - (pt[0, 1] - spec.p[0, 1])
    alpha = 0.9 * gap / (gap - (gap + rel))
    assert it == 1
    np.testing.assert_allclose(sim.p, spec.p + alpha * (pt - spec.p), rtol=0, atol=1e-9)


def test_two_cubes_settle():
    spec = scenes.config_small()
    cp = ContactParams(ground=spec.ground)
    sim = O.OracleSim(spec, cp)
    ys = []
    for _ in range(60):
        sim.step()
        ys.append(sim.p[:, 1].copy())
    ys = np.array(ys)
    gap = ys[:, 1] - ys[:, 0] - 1.0
    assert gap.min() > 0.0                       # never interpenetrates
    assert gap[-1] < 2e-3                        # resting on B within ~dhat
    assert ys[-1, 0] - 0.5 > spec.ground         # B above the ground
    assert np.abs(sim.pdot[:, 1]).max() < 1e-6   # at rest
    assert np.allclose(sim.A, np.eye(3)[None], atol=1e-4)   # nearly rigid (stiffness 1e8)


def test_driver_is_deterministic():
    spec = scenes.config_small()
    cp = ContactParams(ground=spec.ground)
    a, b = O.OracleSim(spec, cp), O.OracleSim(spec, cp)
    for _ in range(15):
        a.step()
        b.step()
    np.testing.assert_array_equal(a.A, b.A)
    np.testing.assert_array_equal(a.p, b.p)


def test_scene_of_splits_a_batch():
    """scene_of() recovers each independent scene of a batch as a one-scene spec
    whose arrays are the batch's slices."""
    spec = scenes.config_batched(n_scenes=3, seed0=5, nb_range=(2, 4))
    assert spec.n_scenes == 3
    b0 = 0
    for s in range(3):
        sub = scenes.scene_of(spec, s)
        nb = sub.n_bodies
        assert sub.n_scenes == 1 and (sub.body_scene == 0).all()
        assert np.array_equal(sub.p, spec.p[b0:b0 + nb])
        v0, v1 = spec.voff[b0], spec.voff[b0 + nb]
        assert np.array_equal(sub.x0, spec.x0[v0:v1])
        assert np.array_equal(sub.mass, spec.mass[b0:b0 + nb])
        b0 += nb
    assert b0 == spec.n_bodies
