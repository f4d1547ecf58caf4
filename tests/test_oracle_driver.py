"""The oracle's frame driver (abd.py:212-276) on the two-cube scene: physical
sanity (A falls onto B and rests at a gap below dhat, B rests on the ground, no
interpenetration) and run-to-run determinism."""
import numpy as np

from abd_hip import scenes
from abd_hip.const_params import ContactParams
from oracle import oracle as O


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
