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
from helpers import driver_pin_cases, free_energy, inertial_target
from oracle import oracle as O

CASES = driver_pin_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_closed_form_frame(name):
    spec, cp, dq, p_want, it_want, tol = CASES[name]
    sim = O.OracleSim(spec, cp)
    if dq is not None:
        sim.dq[...] = dq
    it = sim.step()
    assert it == it_want
    np.testing.assert_allclose(sim.p, p_want, rtol=0, atol=tol)
    if name.startswith("free_fall"):
        np.testing.assert_array_equal(sim.A, np.eye(3)[None])
        np.testing.assert_allclose(sim.pdot, (p_want - spec.p) / cp.dt, rtol=0, atol=1e-10)
        np.testing.assert_array_equal(sim.p0, sim.p)
    if dq is not None:
        np.testing.assert_array_equal(sim.dq, dq)     # the solve kept the planted direction


def test_armijo_case_is_the_as_written_branch():
    """The accepted step of the D8 case raises the energy (E0 < E1 < E0 + c1 q.g), and the
    driver reports that energy as the frame's last."""
    spec, cp, dq, p_want, _, _ = CASES["armijo_as_written_accepts"]
    E0, E1 = free_energy(spec, cp, spec.p[0]), free_energy(spec, cp, p_want[0])
    g = cp.mass * (spec.p[0] - inertial_target(spec, cp)[0])
    assert E0 < E1 < E0 + cp.c1 * float(dq[0] @ g)
    sim = O.OracleSim(spec, cp)
    sim.dq[...] = dq
    sim.step()
    assert sim.E == pytest.approx(E1, rel=1e-12)


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
