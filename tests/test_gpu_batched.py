"""Independent scenes advanced together by the native driver: every scene of a
batch must follow the same Newton path as that scene run alone — against the
oracle's driver (body states to 1e-7 after every frame) on small batches, and
against the plain GPU driver on a sample of the full 256-scene configuration."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _states(sim, spec, s):
    idx = np.nonzero(spec.body_scene == s)[0]
    A = sim.states.A.cpu().numpy()[idx]
    p = sim.states.p.cpu().numpy()[idx]
    return A, p


def test_batched_scenes_match_oracle_per_scene():
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched(n_scenes=5, seed0=7, nb_range=(2, 5))
    cp = ContactParams(ground=spec.ground, solver_tol=1e-12)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    refs = [O.OracleSim(scenes.scene_of(spec, s), cp) for s in range(5)]
    frames = 12
    for f in range(frames):
        sim.step(f)
        its = sim.scene_iters()
        torch.cuda.synchronize()
        for s, ref in enumerate(refs):
            it_ref = ref.step()
            A, p = _states(sim, spec, s)
            err = max(np.abs(A - ref.A).max(), np.abs(p - ref.p).max())
            assert err <= 1e-7, (f, s, err, its[s], it_ref)
    # the scenes fell and touched the floor or each other
    assert sim.stats["gh_active"] > 0


def test_batched_psd_cg_matches_oracle_per_scene():
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched(n_scenes=4, seed0=21, nb_range=(3, 6))
    cp = ContactParams(ground=spec.ground, psd=1, solver=1, solver_tol=1e-10, solver_maxiter=20000)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    refs = [O.OracleSim(scenes.scene_of(spec, s), cp) for s in range(4)]
    for f in range(10):
        sim.step(f)
        torch.cuda.synchronize()
        for s, ref in enumerate(refs):
            ref.step()
            A, p = _states(sim, spec, s)
            err = max(np.abs(A - ref.A).max(), np.abs(p - ref.p).max())
            assert err <= 1e-7, (f, s, err)


def test_batched_iteration_counts_match_plain_gpu():
    """Per-scene Newton iteration counts of the batch equal those of the same scene
    run alone on the GPU, frame by frame."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched(n_scenes=6, seed0=3, nb_range=(2, 8))
    cp = ContactParams(ground=spec.ground, psd=1, solver=1, solver_tol=1e-10, solver_maxiter=20000)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    plain = [AffineBodySimulator(scenes.scene_of(spec, s), cp, "cuda", native=True) for s in range(6)]
    for f in range(6):
        sim.step(f)
        its = sim.scene_iters()
        for s, ps in enumerate(plain):
            it = ps.step(f)
            assert its[s] == it, (f, s, its[s], it)
            A, p = _states(sim, spec, s)
            err = max(np.abs(A - ps.states.A.cpu().numpy()).max(), np.abs(p - ps.states.p.cpu().numpy()).max())
            assert err <= 1e-7, (f, s, err)


def test_full_batched_config_sample_matches_plain():
    """The 256-scene configuration (2..64 bodies per scene): two frames, with three
    sampled scenes checked against the plain driver; every state stays finite."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched()
    cp = ContactParams(ground=spec.ground, psd=1, solver=1)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    sample = [0, 97, 255]
    plain = [AffineBodySimulator(scenes.scene_of(spec, s), cp, "cuda", native=True) for s in sample]
    for f in range(2):
        sim.step(f)
        its = sim.scene_iters()
        assert min(its) >= 1
        for s, ps in zip(sample, plain):
            it = ps.step(f)
            A, p = _states(sim, spec, s)
            err = max(np.abs(A - ps.states.A.cpu().numpy()).max(), np.abs(p - ps.states.p.cpu().numpy()).max())
            assert err <= 1e-7, (f, s, err, its[s], it)
    assert torch.isfinite(sim.states.A).all() and torch.isfinite(sim.states.p).all()
