"""Independent scenes advanced together by the native driver: every scene of a
batch must follow the same Newton path as that scene run alone — against the
oracle's driver (body states to 1e-7 after every frame) on small batches, and
bit for bit against the plain GPU driver, on small batches and on a sample of
the full 256-scene configuration.

Colliding scenes can be ill conditioned: under the reference's indefinite
contact Hessian a 1e-15 relative change of the initial positions grows to O(1)
within a few frames of a collision.  The oracle comparisons therefore run a
second oracle from p (1 + 1e-15) and compare a scene to 1e-7 for as long as
the two oracles agree to 1e-9 (the oracle itself is well conditioned there)."""
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


def _compare_with_oracle(spec, cp, frames):
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    n = spec.n_scenes
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    refs, pert = [], []
    for s in range(n):
        sub = scenes.scene_of(spec, s)
        refs.append(O.OracleSim(sub, cp))
        sub.p = sub.p * (1.0 + 1e-15)
        pert.append(O.OracleSim(sub, cp))
    live = [True] * n
    checked = near_checked = 0
    worst = 0.0
    for f in range(frames):
        sim.step(f)
        its = sim.scene_iters()
        torch.cuda.synchronize()
        for s in range(n):
            c0 = int(refs[s].stats[1] + refs[s].stats[2] + refs[s].stats[3])
            it_ref = refs[s].step()
            pert[s].step()
            near = int(refs[s].stats[1] + refs[s].stats[2] + refs[s].stats[3]) > c0  # gh candidates this frame
            if not live[s]:
                continue
            spread = max(np.abs(refs[s].A - pert[s].A).max(), np.abs(refs[s].p - pert[s].p).max())
            if spread > 1e-9:
                live[s] = False
                continue
            A, p = _states(sim, spec, s)
            err = max(np.abs(A - refs[s].A).max(), np.abs(p - refs[s].p).max())
            assert err <= 1e-7, (f, s, err, its[s], it_ref, spread)
            worst = max(worst, err)
            checked += 1
            near_checked += int(near)
    return checked, near_checked, worst


def test_batched_scenes_match_oracle_per_scene():
    """Reference math (indefinite contact Hessian, BiCGSTAB)."""
    from abd_hip import scenes
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched(n_scenes=5, seed0=7, nb_range=(2, 5))
    cp = ContactParams(ground=spec.ground, solver_tol=1e-12)
    checked, near, worst = _compare_with_oracle(spec, cp, 12)
    # 39 of the 60 (scene, frame) pairs are well conditioned, 27 of them with contact candidates
    assert checked >= 30 and near >= 10, (checked, near)


def test_batched_psd_cg_matches_oracle_per_scene():
    """The stable mode (contact-space PSD projection + Jacobi CG)."""
    from abd_hip import scenes
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched(n_scenes=4, seed0=21, nb_range=(3, 6))
    cp = ContactParams(ground=spec.ground, psd=1, solver=1, solver_tol=1e-10, solver_maxiter=20000)
    checked, near, worst = _compare_with_oracle(spec, cp, 10)
    # 24 of the 40 (scene, frame) pairs are well conditioned, 16 of them with contact candidates
    assert checked >= 20 and near >= 10, (checked, near)


@pytest.mark.parametrize("mode", ["reference", "stable"])
def test_batched_bitwise_equals_plain_gpu(mode):
    """Every reduction is ordered by the scene's own data, so each scene of a batch
    follows exactly the Newton path of the same scene run alone on the plain
    driver: equal iteration counts and bit-identical states, frame by frame, even
    where the path is ill conditioned."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    n = 6
    spec = scenes.config_batched(n_scenes=n, seed0=3, nb_range=(2, 8))
    if mode == "reference":
        cp = ContactParams(ground=spec.ground, solver_tol=1e-12)
    else:
        cp = ContactParams(ground=spec.ground, psd=1, solver=1)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    plain = [AffineBodySimulator(scenes.scene_of(spec, s), cp, "cuda", native=True) for s in range(n)]
    for f in range(8):
        sim.step(f)
        its = sim.scene_iters()
        for s in range(n):
            it = plain[s].step(f)
            assert its[s] == it, (f, s, its[s], it)
            A, p = _states(sim, spec, s)
            assert np.array_equal(A, plain[s].states.A.cpu().numpy()), (f, s)
            assert np.array_equal(p, plain[s].states.p.cpu().numpy()), (f, s)


def test_full_batched_config_sample_matches_plain():
    """The 256-scene configuration (2..64 bodies per scene) in the benchmark mode:
    three frames, with sampled scenes bit-identical to the plain driver; every
    state stays finite."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched()
    cp = ContactParams(ground=spec.ground, psd=1, solver=1)
    sim = AffineBodySimulator(spec, cp, "cuda", native=True)
    sample = [0, 97, 255]
    plain = [AffineBodySimulator(scenes.scene_of(spec, s), cp, "cuda", native=True) for s in sample]
    for f in range(3):
        sim.step(f)
        its = sim.scene_iters()
        assert min(its) >= 1
        for s, ps in zip(sample, plain):
            it = ps.step(f)
            assert its[s] == it, (f, s, its[s], it)
            A, p = _states(sim, spec, s)
            assert np.array_equal(A, ps.states.A.cpu().numpy()), (f, s)
            assert np.array_equal(p, ps.states.p.cpu().numpy()), (f, s)
    assert torch.isfinite(sim.states.A).all() and torch.isfinite(sim.states.p).all()


def test_autotuned_tiles_change_no_result():
    """Different launch shapes (the culling tiles, stage D's grid: 64 workgroups make each sum
    several items, and the CCD root searches' grid; the autotuner's choices) give
    bit-identical frames."""
    from abd_hip import _lib, scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    import ctypes
    spec = scenes.config_batched(n_scenes=8, seed0=11, nb_range=(2, 12))
    cp = ContactParams(ground=spec.ground, psd=1, solver=1)
    L = _lib.load()
    t0 = _lib.AbdTuning()
    L.abd_get_tuning(ctypes.byref(t0))
    out = []
    try:
        for chunk, grid, pgrid, tgrid in ((4096, 16384, 0, 0), (1024, 64, 64, 64), (65536, 4096, 8192, 4096)):
            assert L.abd_set_tuning(ctypes.byref(_lib.AbdTuning(chunk, grid, 0, pgrid, tgrid))) == 0
            sim = AffineBodySimulator(spec, cp, "cuda", native=True)
            for f in range(5):
                sim.step(f)
            out.append((sim.states.A.cpu().numpy(), sim.states.p.cpu().numpy(), dict(sim.stats)))
        chosen = sim.autotune()
        assert chosen["pair_grid"] in (1024, 2048), chosen
    finally:
        L.abd_set_tuning(ctypes.byref(t0))
    for A, p, st in out[1:]:
        assert np.array_equal(A, out[0][0]) and np.array_equal(p, out[0][1]) and st == out[0][2]


def test_contact_bound_read_back_equals_bound_sizing():
    """Sizing the contact buffer by the active count read back (taken when the candidate
    lists' capacity would need more than contact_bound_mib) gives the states, iteration
    counts and statistics of sizing it by that capacity; those iterations are launched, not
    replayed from a graph."""
    import ctypes
    from abd_hip import _lib, scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    L = _lib.load()
    t0 = _lib.AbdTuning()
    assert L.abd_get_tuning(ctypes.byref(t0)) == 0
    spec = scenes.config_batched(n_scenes=16, seed0=41, nb_range=(2, 24))
    cp = ContactParams(ground=spec.ground, psd=1, solver=3)
    sized = AffineBodySimulator(spec, cp, "cuda", native=True)
    for f in range(10):
        sized.step(f)
    torch.cuda.synchronize()
    try:
        # 1 MiB: every iteration with more than about a thousand candidates reads back
        assert L.abd_set_tuning(ctypes.byref(_lib.AbdTuning(t0.cull_chunk, t0.cull_grid, 1))) == 0
        read = AffineBodySimulator(spec, cp, "cuda", native=True)
        for f in range(10):
            read.step(f)
        torch.cuda.synchronize()
    finally:
        assert L.abd_set_tuning(ctypes.byref(t0)) == 0
    assert sized.scene_iters() == read.scene_iters()
    assert np.array_equal(sized.states.A.cpu().numpy(), read.states.A.cpu().numpy())
    assert np.array_equal(sized.states.p.cpu().numpy(), read.states.p.cpu().numpy())
    assert dict(sized.stats) == dict(read.stats)
    assert read.graph_counts()["replayed"] < sized.graph_counts()["replayed"], (read.graph_counts(),
                                                                               sized.graph_counts())


def test_step_end_projection_unaligned_equals_pairs():
    """The step's end projected by vertex pairs with 16-byte accesses (vertex buffers 16-byte
    aligned) and by the per-vertex kernel (the buffers 8 bytes off) gives the same frames."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.scene_of(scenes.config_batched(n_scenes=4, seed0=5, nb_range=(4, 8)), 2)
    cp = ContactParams(ground=spec.ground, psd=1, solver=3)
    aligned = AffineBodySimulator(spec, cp, "cuda", native=True)
    shifted = AffineBodySimulator(spec, cp, "cuda", native=False)
    m = shifted.meshes
    for name in ("x", "xk"):
        t = getattr(m, name)
        buf = torch.empty(t.numel() + 1, dtype=t.dtype, device=t.device)
        view = buf[1:].view_as(t)
        view.copy_(t)
        assert view.data_ptr() % 16 == 8
        setattr(m, name, view)
    shifted._make_native()
    for f in range(6):
        ia = aligned.step(f)
        ib = shifted.step(f)
        assert ia == ib, (f, ia, ib)
        assert np.array_equal(aligned.states.A.cpu().numpy(), shifted.states.A.cpu().numpy()), f
        assert np.array_equal(aligned.states.p.cpu().numpy(), shifted.states.p.cpu().numpy()), f


@pytest.mark.parametrize("solver", [3, 1])
def test_graph_replay_equals_launches(solver):
    """Newton iterations replayed from a recorded HIP graph give the states, iteration
    counts and statistics of the same iterations launched kernel by kernel, frame after
    frame (each frame's replays run between eager work and host waits), and the replay
    path is taken."""
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    spec = scenes.config_batched(n_scenes=16, seed0=41, nb_range=(2, 24))
    cp = ContactParams(ground=spec.ground, psd=1, solver=solver)
    eager = AffineBodySimulator(spec, cp, "cuda", native=True)
    eager.set_graphs(False)
    graph = AffineBodySimulator(spec, cp, "cuda", native=True)
    for f in range(14):
        eager.step(f)
        graph.step(f)
        torch.cuda.synchronize()
        assert eager.scene_iters() == graph.scene_iters(), f
        assert np.array_equal(eager.states.A.cpu().numpy(), graph.states.A.cpu().numpy()), f
        assert np.array_equal(eager.states.p.cpu().numpy(), graph.states.p.cpu().numpy()), f
        assert dict(eager.stats) == dict(graph.stats), f
    gc = graph.graph_counts()
    assert gc["replayed"] >= 10 and gc["recorded"] >= 1, gc
    assert eager.graph_counts()["replayed"] == 0
