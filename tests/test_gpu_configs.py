"""The benchmark's mode and the full-size medium configuration against the oracle,
and a long property run of the medium configuration.

* The 256-scene batch in the benchmark mode (contact-space PSD projection, direct
  block LDL^T solve): scenes 0, 97 and 255 (14, 7 and 36 bodies) against the
  oracle's driver run on each scene alone, 12 frames with contact: each frame from
  the GPU's state to 1e-7 with equal Newton iteration counts, and the free-running
  trajectories to 1e-7 or the scene's own rounding sensitivity.
* The medium configuration at full size (64 icosphere-3 bodies, 41k vertices) in
  the same mode: 12 frames against the oracle (contacts from frame 5 on), then
  200 frames on the GPU alone with the separation checked every 50 frames: finite
  states, no edge crossing a triangle of another body, every point-triangle,
  edge-edge and vertex-ground distance positive, and bodies resting on the floor
  within dhat of it (tests/geometry.py)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _scene_states(sim, spec, s):
    idx = np.nonzero(spec.body_scene == s)[0]
    return sim.states.A.cpu().numpy()[idx], sim.states.p.cpu().numpy()[idx]


def _seed(ref, sim, spec, s):
    """The oracle's frame-start state (q0, qdot, warm start) set to the GPU's for scene s."""
    idx = np.nonzero(spec.body_scene == s)[0]
    st = sim.states
    for name in ("A0", "p0", "Adot", "pdot"):
        getattr(ref, name)[...] = getattr(st, name).cpu().numpy()[idx]
    dq = sim.dq.cpu().numpy().reshape(-1, 4, 3)[idx].reshape(-1, 3)
    ref.dq[...] = dq


def test_benchmark_mode_sampled_scenes_match_oracle():
    """Two comparisons per frame.  Step parity: the oracle started from the GPU's state
    at the start of the frame ends it within 1e-7 of the GPU with the same Newton
    iteration count.  Trajectory parity: an oracle run free from frame 0 stays within
    1e-7 of the GPU, or within 10 times the distance between it and a second free oracle
    whose initial positions differ by a relative 1e-15 where that distance is larger:
    these colliding stacks, capped at 10 Newton iterations per frame, amplify rounding
    (a 1e-15 change grows to 1e-5 within 6 frames in scene 0), so no two FP64 evaluation
    orders can stay closer than the scene's own sensitivity."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched()
    cp = ContactParams(ground=spec.ground, psd=1, solver=3)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    sample = [0, 97, 255]
    step_refs = [O.OracleSim(scenes.scene_of(spec, s), cp) for s in sample]
    free = [O.OracleSim(scenes.scene_of(spec, s), cp) for s in sample]
    pert = []
    for s in sample:
        sub = scenes.scene_of(spec, s)
        sub.p = sub.p * (1.0 + 1e-15)
        pert.append(O.OracleSim(sub, cp))
    worst_step = 0.0
    for f in range(12):
        for s, ref in zip(sample, step_refs):
            _seed(ref, sim, spec, s)
        sim.step(f)
        its = sim.scene_iters()
        for k, s in enumerate(sample):
            it_ref = step_refs[k].step()
            free[k].step()
            pert[k].step()
            A, p = _scene_states(sim, spec, s)
            err = max(np.abs(A - step_refs[k].A).max(), np.abs(p - step_refs[k].p).max())
            err_free = max(np.abs(A - free[k].A).max(), np.abs(p - free[k].p).max())
            spread = max(np.abs(free[k].A - pert[k].A).max(), np.abs(free[k].p - pert[k].p).max())
            print(f"frame {f} scene {s}: newton {its[s]} / oracle {it_ref}, step error {err:.2e}, "
                  f"trajectory error {err_free:.2e}, oracle spread {spread:.2e}", flush=True)
            assert its[s] == it_ref, (f, s)
            assert err <= 1e-7, (f, s, err)
            assert err_free <= max(1e-7, 10 * spread), (f, s, err_free, spread)
            worst_step = max(worst_step, err)
    for ref in step_refs:
        assert ref.stats[1] + ref.stats[2] + ref.stats[3] > 0   # the frames had contact candidates
    print(f"worst step error {worst_step:.2e}")


def _positions(spec, A, p):
    b = np.repeat(np.arange(spec.n_bodies), np.diff(spec.voff))
    return np.einsum("vji,vj->vi", A[b], spec.x0) + p[b]


@pytest.fixture(scope="module")
def medium_run():
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_medium()
    cp = ContactParams(ground=spec.ground, psd=1, solver=3)
    return spec, cp, AffineBodySimulator(spec, cp, "cuda", native=True)


def test_medium_full_size_matches_oracle(medium_run):
    spec, cp, sim = medium_run
    ref = O.OracleSim(spec, cp)
    for f in range(12):
        it = sim.step(f)
        it_ref = ref.step()
        A, p = sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()
        err = max(np.abs(A - ref.A).max(), np.abs(p - ref.p).max())
        print(f"frame {f}: newton {it} / oracle {it_ref}, state error {err:.2e}", flush=True)
        assert it == it_ref, f
        assert err <= 1e-7, (f, err)
    assert ref.stats[1] + ref.stats[2] > 1000     # contact candidates in the Newton iterations
    assert sim.stats["gh_active"] > 0


def test_medium_full_size_200_frames_stay_separated(medium_run):
    import geometry as G
    spec, cp, sim = medium_run
    f0 = sim.stats["frames"]
    for f in range(f0, 200):
        sim.step(f)
        if (f + 1) % 50 == 0:
            A, p = sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()
            assert np.isfinite(A).all() and np.isfinite(p).all(), f
            r = G.separation(spec, _positions(spec, A, p), spec.ground)
            print(f"frame {f}: {r}", flush=True)
            assert r["crossings"] == 0, (f, r)
            assert min(r["min_pt"], r["min_ee"]) > 0 and r["min_vg"] > 0, (f, r)
    # bodies came to rest on the floor: the lowest vertex sits inside the barrier's range
    assert 0 < r["min_vg"] < cp.dhat, r


@pytest.mark.parametrize("which", ["stress", "stress_batched"])
def test_stress_300_frames_stay_separated(which):
    """The large stress configuration at full size (1024 bodies: cubes, icosphere-3 and
    icosphere-4 ellipsoids, 1.1M vertices, packed with gap 0.02) and its batched form
    (8 scenes of 128 bodies), 300 frames in the benchmark mode: the pile collapses onto
    the floor with thousands of contact candidates per Newton iteration by the end;
    states stay finite, no edge crosses a triangle of another body and every distance
    stays positive (checked at frames 100, 200 and 300)."""
    import geometry as G
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_stress() if which == "stress" else scenes.config_stress_batched()
    cp = ContactParams(ground=spec.ground, psd=1, solver=3)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    for f in range(300):
        s0 = dict(sim.stats)
        sim.step(f)
        if (f + 1) % 100 == 0:
            torch.cuda.synchronize()
            n_it = sim.stats["newton_iters"] - s0["newton_iters"]
            cands = (sim.stats["gh_pt"] + sim.stats["gh_ee"] - s0["gh_pt"] - s0["gh_ee"]) / max(n_it, 1)
            A, p = sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy()
            assert np.isfinite(A).all() and np.isfinite(p).all(), f
            r = G.separation(spec, _positions(spec, A, p), spec.ground)
            print(f"frame {f}: {cands:.0f} PT + EE candidates per Newton iteration; {r}", flush=True)
            assert r["crossings"] == 0, (f, r)
            assert min(r["min_pt"], r["min_ee"]) > 0 and r["min_vg"] > 0, (f, r)
    assert cands > 1000 and sim.stats["gh_active"] > 0
