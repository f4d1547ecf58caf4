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


def test_scene_of_splits_a_batch():
    """scene_of() recovers each independent scene of a batch as a one-scene spec
    whose arrays are the batch's slices."""
    from abd_hip import scenes
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
