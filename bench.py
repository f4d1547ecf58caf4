"""Benchmark of the per-Newton-iteration contact path on one MI355X.

Workload (default, "batched"): 256 independent scenes advanced together, scene
s seeded 1000 + s; each draws U{2..64} bodies, each a cube (8 V), icosphere-2
(162 V) or 6x6x6 subdivided box (218 V), stacked in random cells of a 4 x 4
grid with gap 0.02 over a floor, speed U[0, 2] in a uniform direction
(8800 bodies, 1.13M vertices, 2.22M triangles in all), dhat 1e-3, kappa 1e7,
orthogonality stiffness 1e8, dt 0.01, FP64.  A "step" is one frame of every
scene (AffineBodySimulator.step: per scene, Newton iterations of culling,
barrier gradient/Hessian, solve, CCD and line search; converged scenes are
masked).  `value` is the sum over scenes of Newton iterations per second of
wall time.  Beside it: the plain path (one scene per simulator, the same
kernels) timed on 8 sampled scenes over the same frames, the ratio of the
two, and the gh-stage throughput (barrier-Hessian candidate pairs per second
of gh kernel time).  "--workload medium" runs the 64-icosphere plain scene.

Solver mode: with the reference's contact Hessian (indefinite: b' < 0 times a
clamped-PSD matrix) the reference's BiCGSTAB diverges on colliding scenes and
the swept CCD boxes then cover the scene.  The benchmark therefore runs the
documented stable mode: contact-space PSD projection of the barrier Hessian
(psd=1), with each Newton system solved directly by a block LDL^T factorization
(solver 3; systems above 64 bodies fall back to the block-Jacobi CG).
--solver 2 runs the reference's CG (bsr_utils.py:126-235, same iteration and
stopping rule) preconditioned by each body's 12x12 diagonal block, --solver 1
with its scalar diagonal.  The CPU baseline runs the oracle in the same mode on
the sampled scenes (1 core) and, over every host core, on the whole batch's scenes
dealt out in turn.

Frames: W warmup frames, then (unless --no-profile) untimed attribution frames in
which every instrumented kernel is timed with HIP events (the per-kernel table,
the gh-stage rate and the choice of the dominant kernel), then the K timed
frames with no kernel timing (`value`; the batched driver replays its recorded
Newton iterations as HIP graphs there), then --roof-frames frames (default K) in
which only the dominant kernel carries events (its roofline; with kernel timing on
every kernel is launched on its own).

"--workload stress" runs the 1024-body lattice (and "stress_batched" 8 scenes of 128
bodies); their contact phase starts once the pile has formed, so run them with
"--warmup 200" or so.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
                       [--workload batched|medium|small|stress|stress_batched]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "abd-warp_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str):
    from abd_hip import scenes
    if name == "medium":
        return scenes.config_medium(), "medium: 64 icosphere-3 bodies (642 V / 1280 F / 1920 E), 4x4x4 grid, floor"
    if name == "small":
        return scenes.config_small(), "small: 2 unit cubes (8 V / 12 F / 18 E), B at origin, A at 1.05"
    if name == "stress":
        spec = scenes.config_stress()
        return spec, (f"stress: 1024 bodies packed in a jittered 10x10x10.24 lattice, gap 0.02 (cube / icosphere-3 "
                      f"/ perturbed icosphere-4 ellipsoid in equal thirds; {spec.n_verts} V / {spec.n_tris} F), "
                      "collapsing onto a floor")
    if name == "stress_batched":
        spec = scenes.config_stress_batched()
        return spec, (f"stress_batched: 8 independent scenes of 128 bodies of the stress mix, each packed in an "
                      f"8x8x2 lattice with gap 0.02 ({spec.n_verts} V / {spec.n_tris} F in all), collapsing onto a "
                      "floor")
    if name == "batched":
        spec = scenes.config_batched()
        return spec, (f"batched: 256 independent scenes, U{{2..64}} bodies each ({spec.n_bodies} bodies, "
                      f"{spec.n_verts} V / {spec.n_tris} F in all; cube, icosphere-2 or 6x6x6 box), stacked, "
                      "speed U[0,2], floor")
    raise SystemExit(f"unknown workload {name}")


GH_KERNELS = ("gh_classify", "gh_contact", "gh_pair_reduce", "gh_body_finalize")
# measured HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a copy of known
# size) of the profiled kernels, kept under profiles/ as "<kernel>_<workload>_pmc.md"
PMC_FILES = {"ccd_toi": "toi_scan_kernel", "solve": "ldlt_kernel", "gh_pair_reduce": "pair_partial_rank_kernel",
             "energy": "chunk_kernel_energy"}


def measured_traffic(kernel: str, workload: str, solver: int = 3):
    name = PMC_FILES.get(kernel)
    if kernel == "solve" and solver != 3:
        name = None  # the Krylov kernels' traffic is not profiled for this configuration
    if not name:
        return None
    path = os.path.join(ROOT, "profiles", f"{name}_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def plain_per_scene(spec, cp, sample, warmup: int, steps: int):
    """The plain (one scene per simulator) GPU path on sampled scenes of a batch:
    Newton iterations per second of wall time over the same frames as the timed
    batched region."""
    import torch
    from abd_hip import scenes
    from abd_hip.abd import AffineBodySimulator
    iters, wall = 0, 0.0
    for s in sample:
        sim = AffineBodySimulator(scenes.scene_of(spec, s), cp, "cuda", native=True)
        for f in range(warmup):
            sim.step(f)
        torch.cuda.synchronize()
        t = time.perf_counter()
        n = 0
        for f in range(steps):
            n += sim.step(warmup + f)
        torch.cuda.synchronize()
        wall += time.perf_counter() - t
        iters += n
        log(f"[plain] scene {s}: {n} Newton iterations")
        del sim
    return iters, wall


def reference_solver_line(spec, psd: int, first: int, steps: int):
    """The batch with the reference's solver, the Jacobi-preconditioned CG of
    bsr_utils.py:126-235 at its tolerance 1e-4 (solver 1), on a simulator of its own
    over the same frames as the timed region: Newton iterations per second."""
    import torch
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams
    cp1 = ContactParams(ground=spec.ground, psd=psd, solver=1)
    sim = AffineBodySimulator(spec, cp1, "cuda", native=True)
    for f in range(first):
        sim.step(f)
    torch.cuda.synchronize()
    n0 = sim.stats["newton_iters"]
    t = time.perf_counter()
    for f in range(first, first + steps):
        sim.step(f)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    n = sim.stats["newton_iters"] - n0
    log(f"[gpu] reference solver (Jacobi CG): {n} Newton iterations in {wall:.3f} s")
    return dict(value=n / wall, unit="Newton iters/s", solver="jacobi-cg", ms_per_step=1e3 * wall / steps,
                sample=f"batched native driver, solver 1 (Jacobi CG, tol 1e-4), psd {psd}, frames "
                       f"{first}..{first + steps - 1}: {n} Newton iterations in {wall:.3f} s")


def gpu_state(sim, bodies):
    """Frame-start state (A0, p0, Adot, pdot, warm start dq) of some bodies of a GPU
    simulator, on the host."""
    import numpy as np
    st = sim.states
    out = {k: getattr(st, k).cpu().numpy()[bodies] for k in ("A0", "p0", "Adot", "pdot")}
    out["dq"] = sim.dq.cpu().numpy().reshape(-1, 4, 3)[bodies].reshape(-1, 3)
    return out


def cpu_baseline(spec, cp, warmup: int, steps: int, budget_s: float, sample=None, start=None):
    """The oracle (C FP64 restatement, 1 thread) on the same frames as the timed
    GPU region, stopped after `budget_s` seconds of timed CPU work.  With `sample`
    (scene ids of a batch), the sampled scenes run one after another, each from
    its own frame 0, sharing the budget.  With `start` (one state per sampled
    scene, gpu_state), each starts from the GPU's state at the first timed frame
    instead of simulating the warmup frames."""
    from oracle import oracle as O
    from abd_hip import scenes
    specs = [spec] if sample is None else [scenes.scene_of(spec, s) for s in sample]
    iters, t_cpu, frames, parts, frames_done = 0, 0.0, 0, [], []
    share = budget_s / len(specs)
    for k, sp in enumerate(specs):
        ref = O.OracleSim(sp, cp)
        if start is not None:
            for name, v in start[k].items():
                getattr(ref, name)[...] = v
        for _ in range(warmup if start is None else 0):
            ref.step()
        it0 = int(ref.stats[0])
        t_here, f_here = 0.0, 0
        for _ in range(steps):
            t = time.perf_counter()
            ref.step()
            t_here += time.perf_counter() - t
            f_here += 1
            if t_here >= share:
                break
        n = int(ref.stats[0]) - it0
        log(f"[cpu] part {k}: {f_here} frames, {n} Newton iterations, {t_here:.2f}s")
        iters += n
        t_cpu += t_here
        frames += f_here
        frames_done.append(f_here)
        parts.append(f"{'scene ' + str(sample[k]) + ': ' if sample is not None else ''}frames {warmup}.."
                     f"{warmup + f_here - 1}, {n} it")
    return dict(value=iters / t_cpu if t_cpu > 0 else None, unit="Newton iters/s", cores=1, kind="port",
                sample=f"oracle (C FP64 restatement, -O2, 1 thread), same mode: " + "; ".join(parts) +
                       f"; {iters} Newton iterations in {t_cpu:.2f} s"), frames_done


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def host_cores() -> int:
    """Cores this process may run on (the box's share, not the machine's count)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_worker(subs, cp, warmup, frames, budget, barrier, q, k):
    # one host core: its share of the batch's scenes on the oracle, one scene after another,
    # each advanced through the untimed warmup frames and then the timed frames; the timed
    # frames' seconds are summed, and the worker stops after the scene that uses up the budget
    from oracle import oracle as O
    barrier.wait()
    iters, timed, fdone, nsc = 0, 0.0, 0, 0
    for sp in subs:
        ref = O.OracleSim(sp, cp)
        for _ in range(warmup):
            ref.step()
        it0 = int(ref.stats[0])
        t = time.perf_counter()
        for _ in range(frames):
            ref.step()
        timed += time.perf_counter() - t
        iters += int(ref.stats[0]) - it0
        fdone += frames
        nsc += 1
        if timed >= budget:
            break
    q.put((k, iters, timed, fdone, nsc))


def cpu_baseline_parallel(spec, cp, warmup: int, frames: int, budget_s: float, n_cores: int):
    """All host cores on the batch: one oracle process per core, the batch's scenes dealt
    out in turn (core k takes scenes k, k + n, ...), each scene run through the same frames
    as the timed GPU region after its warmup frames; a core stops after the scene that
    takes its timed seconds past `budget_s`.  Each core's rate (its Newton iterations per
    second of its timed frames, over a mix of small and large scenes) is summed over the
    cores: the throughput of all cores working through the batch."""
    import multiprocessing as mp
    from abd_hip import scenes
    ctx = mp.get_context("spawn")  # fresh interpreters: nothing of this process's GPU state
    ns = spec.n_scenes
    n = max(1, min(n_cores, ns))
    barrier, q = ctx.Barrier(n), ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker, args=([scenes.scene_of(spec, s) for s in range(k, ns, n)], cp, warmup,
                                                   frames, budget_s, barrier, q, k))
             for k in range(n)]
    import queue
    for p in procs:
        p.start()
    res = []
    try:
        for _ in procs:
            res.append(q.get(timeout=900))
    except queue.Empty:
        for p in procs:
            p.terminate()
        log("[cpu] parallel baseline: a worker did not finish")
        return None
    for p in procs:
        p.join()
    iters = sum(r[1] for r in res)
    rate = sum(r[1] / r[2] for r in res if r[2] > 0)
    fdone = sum(r[3] for r in res)
    nsc = sum(r[4] for r in res)
    return dict(value=rate, unit="Newton iters/s", cores=n, kind="port", cpu_model=cpu_model(),
                sample=f"oracle (C FP64 restatement, -O2), one process per core, the batch's scenes dealt out in "
                       f"turn: {nsc} of {ns} scenes run, frames {warmup}..{warmup + frames - 1} each ({fdone} "
                       f"frames, {iters} Newton iterations, about {budget_s:.0f} s of timed frames per core); the "
                       f"sum over cores of iterations per second of timed frames")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batched")
    ap.add_argument("--psd", type=int, default=1)
    ap.add_argument("--solver", type=int, default=3,
                    help="0 BiCGSTAB, 1 Jacobi CG, 2 block-Jacobi CG, 3 direct block LDL^T")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=8,
                    help="batched: scenes timed on the plain path (0: none)")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-kernel HIP events (no attribution frames, roofline or kernel breakdown)")
    ap.add_argument("--attrib-frames", type=int, default=2,
                    help="untimed frames with every instrumented kernel timed, before the timed frames")
    ap.add_argument("--no-graphs", dest="graphs", action="store_false",
                    help="batched native driver: launch every kernel (no HIP graph replay of Newton iterations)")
    ap.add_argument("--roof-frames", type=int, default=None,
                    help="frames after the timed ones with HIP events on the dominant kernel (default: --steps)")
    ap.add_argument("--roof-stages", default="",
                    help="comma-separated profiled stages also timed in the roofline frames (e.g. gh_pair_reduce)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with one 256 MiB device copy (reference kernel for HBM counter scaling)")
    ap.add_argument("--python-driver", action="store_true", help="run the Python mirror of the driver")
    ap.add_argument("--no-ref-solver", dest="ref_solver", action="store_false",
                    help="batched: skip the line with the reference's Jacobi CG beside the chosen solver")
    a = ap.parse_args()
    if a.roof_frames is None:
        a.roof_frames = a.steps
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the contact path of one scene is "
              "not split across GPUs in this build", file=sys.stderr)
        sys.exit(2)

    import torch
    import abd_hip
    from abd_hip import _lib
    from abd_hip.abd import AffineBodySimulator
    from abd_hip.const_params import ContactParams

    abd_hip.load()
    spec, desc = make_workload(a.workload)
    cp = ContactParams(ground=spec.ground, psd=a.psd, solver=a.solver)
    sim = AffineBodySimulator(spec, cp, "cuda", native=not a.python_driver)
    if not a.python_driver and not a.graphs:
        sim.set_graphs(False)
    for f in range(a.warmup):
        it = sim.step(f)
        log(f"[gpu] warmup frame {f} newton {it}")
    tiles = None
    if not a.python_driver:
        # launch shapes of the culling chunk passes, timed at the warmed-up state (no effect on results)
        tiles = sim.autotune()
        log(f"[gpu] autotuned launch shapes: {tiles}")
    frame = a.warmup
    # attribution frames (untimed): every instrumented kernel timed with HIP events, for the
    # per-kernel table, the gh-stage rate and the candidates for the dominant kernel; the timed
    # frames then carry events on those only, so the event overhead stays out of `value`
    aprof, dom = None, None
    gh_s, gh_cands = 0.0, 0
    n_attrib = 0 if a.no_profile else max(a.attrib_frames, 1)
    if n_attrib:
        torch.cuda.synchronize()
        sa = dict(sim.stats)
        _lib.profile_select(None)
        _lib.profile_begin()
        for f in range(n_attrib):
            it = sim.step(frame)
            log(f"[gpu] attribution frame {frame} newton {it}")
            frame += 1
        torch.cuda.synchronize()
        _lib.profile_end()
        aprof = _lib.profile_query()
        # gh rate over these frames (both drivers launch the same four gh kernels)
        gh_s = sum(aprof[k][0] for k in GH_KERNELS if k in aprof) * 1e-3
        gh_cands = sim.stats["gh_pt"] + sim.stats["gh_ee"] - sa["gh_pt"] - sa["gh_ee"]
        # the dominant kernel: the largest with an algorithmic byte count
        ranked = sorted(((v[0], k) for k, v in aprof.items() if v[1] > 0 and v[2] > 0), reverse=True)
        # the two largest: the first is the dominant kernel; its runner-up gets a roofline too
        # (their order can change from one run to the next), and so does the solve
        cands_dom = [k for _, k in ranked[:2]]
        if "solve" in aprof and "solve" not in cands_dom and aprof["solve"][2] > 0:
            cands_dom.append("solve")
        for k in filter(None, a.roof_stages.split(",")):
            if k not in _lib.PROF_NAMES:
                raise SystemExit(f"unknown stage {k}: one of {_lib.PROF_NAMES}")
            if k not in cands_dom and k in aprof and aprof[k][1] > 0:
                cands_dom.append(k)
        dom = cands_dom[0] if cands_dom else None
    first = frame
    torch.cuda.synchronize()
    start_state = None
    if a.workload == "stress":
        start_state = gpu_state(sim, [i + 10 * j + 100 * k for k in range(3) for j in range(3) for i in range(3)])
    elif a.workload == "stress_batched":
        import numpy as np
        start_state = gpu_state(sim, np.nonzero(spec.body_scene == 0)[0])
    torch.cuda.synchronize()
    s0 = dict(sim.stats)
    g0 = sim.graph_counts() if not a.python_driver else None
    # the timed frames: no kernel timing, so the batched driver replays its recorded Newton
    # iterations (HIP graphs) as it does in production
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(a.steps):
        it = sim.step(first + f)
        log(f"[gpu] frame {first + f} newton {it} elapsed {time.perf_counter() - t0:.2f}s")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    d = {k: sim.stats[k] - s0[k] for k in sim.stats}
    graphs = None
    if g0 is not None:
        g1 = sim.graph_counts()
        graphs = {k: g1[k] - g0[k] for k in g1}
    prof = None
    roof_first = first + a.steps
    if dom:
        # roofline frames, after the timed ones: HIP events on the dominant kernel's launches
        # only (with kernel timing on, every kernel is launched on its own, no graph replay)
        _lib.profile_select(sorted(cands_dom))
        _lib.profile_begin()
        for f in range(a.roof_frames):
            it = sim.step(roof_first + f)
            log(f"[gpu] roofline frame {roof_first + f} newton {it}")
        torch.cuda.synchronize()
        _lib.profile_end()
        prof = _lib.profile_query()
        _lib.profile_select(None)
    if a.calibrate_copy:
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[gpu] calibration copy: 268435456 bytes read, 268435456 bytes written")
        del src, dst
    wall = t1 - t0
    iters = d["newton_iters"]
    cands = d["gh_pt"] + d["gh_ee"]
    roof = None
    rooflines = None
    timed_kernels = None
    if dom:
        timed_kernels = {k: dict(launches=prof[k][1], avg_us=round(1e3 * prof[k][0] / prof[k][1], 3),
                                 ms_per_frame=round(prof[k][0] / a.roof_frames, 3))
                         for k in cands_dom if k in prof and prof[k][1] > 0}
        dom = max(timed_kernels, key=lambda k: timed_kernels[k]["ms_per_frame"])

        def roofline_of(k):
            ms, n, byts = prof[k]
            achieved = byts / (ms * 1e-3) / 1e9
            return dict(kernel=k, bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=achieved / HBM_PEAK_GBS, traffic=measured_traffic(k, a.workload, a.solver), launches=n,
                        avg_us=round(1e3 * ms / n, 3), bytes_per_launch=byts / n)
        roof = roofline_of(dom)
        rooflines = {k: roofline_of(k) for k in timed_kernels}
    batched = spec.n_scenes > 1
    plain = None
    sample = None
    if batched and a.plain_sample > 0:
        ns = spec.n_scenes
        k_n = a.plain_sample
        sample = [0] if k_n == 1 else sorted(set(int(round(k * (ns - 1) / (k_n - 1))) for k in range(k_n)))
        p_it, p_wall = plain_per_scene(spec, cp, sample, first, a.steps)
        plain = dict(value=p_it / p_wall if p_wall > 0 else None, unit="Newton iters/s",
                     sample=f"plain native driver, one scene at a time, scenes {sample}, frames "
                            f"{first}..{first + a.steps - 1}: {p_it} Newton iterations in {p_wall:.3f} s")
    ref_solver = None
    if a.workload == "batched" and a.solver != 1 and a.ref_solver:
        ref_solver = reference_solver_line(spec, a.psd, first, a.steps)
    cpu, cpu_par = None, None
    if not a.no_cpu:
        if a.workload == "stress":
            # the oracle's brute-force culling cannot run the full stress scene in a bounded time:
            # its sample is the 3 x 3 x 3 corner block of the lattice (floor layer included) from
            # the GPU's state at the first timed frame, same frames
            from abd_hip import scenes
            corner = [i + 10 * j + 100 * k for k in range(3) for j in range(3) for i in range(3)]
            sub = scenes.subset_scene(spec, corner)
            cpu, _ = cpu_baseline(sub, cp, first, a.steps, a.cpu_budget, start=[start_state])
            cpu["sample"] = ("27 bodies (the 3x3x3 corner of the lattice) from the GPU's state at frame "
                             f"{first}, " + cpu["sample"])
        elif a.workload == "stress_batched":
            # one scene of 128 bodies from the GPU's state at the first timed frame
            cpu, _ = cpu_baseline(spec, cp, first, a.steps, a.cpu_budget, [0], start=[start_state])
            cpu["sample"] = f"scene 0 from the GPU's state at frame {first}, " + cpu["sample"]
        else:
            cpu_sample = sample if sample is not None or not batched else [0]
            cpu, frames_done = cpu_baseline(spec, cp, first, a.steps, a.cpu_budget, cpu_sample)
            if a.workload == "batched":
                # every host core on its own scene of the batch (scenes are independent)
                cpu_par = cpu_baseline_parallel(spec, cp, first, a.steps, a.cpu_budget, host_cores())
        cpu["cpu_model"] = cpu_model()
    out = {
        "metric": "Newton iters/s + barrier-Hessian pairs/s, batched vs plain, 1 MI355X",
        "value": iters / wall,
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": desc, "mode": "batched" if batched else "plain", "psd_projection": bool(a.psd),
                   "driver": "python" if a.python_driver else "native",
                   "solver": {0: "bicgstab", 1: "jacobi-cg", 2: "block-jacobi-cg", 3: "block-ldlt"}[a.solver],
                   "frames_timed": f"{first}..{first + a.steps - 1}", "tiles": tiles,
                   "frames_roofline": (f"{roof_first}..{roof_first + a.roof_frames - 1}" if dom else None),
                   "graph_iterations": graphs},
        # PT + EE candidates through the four gh kernels per second of their event time
        # (attribution frames)
        "gh_pairs_per_s": gh_cands / gh_s if gh_s > 0 and gh_cands > 0 else None,
        "gh_candidates": cands,
        "gh_active": d["gh_active"],
        "newton_iters": iters,
        "ccd_candidates": d["ccd_pt"] + d["ccd_ee"],
        "plain_per_scene": plain,
        "batched_over_plain": (iters / wall) / plain["value"] if plain and plain["value"] else None,
        "roofline": roof,
        "rooflines": rooflines,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_par,
        "host_cores": host_cores(),
        # the same batch with the reference's solver (bsr_utils.py:126-235 Jacobi CG), same frames
        "reference_solver": ref_solver,
        # per-frame kernel times of the attribution frames (all instrumented kernels timed there;
        # the timed frames time the dominant kernel only)
        "kernels_ms_per_frame": {k: round(v[0] / n_attrib, 3) for k, v in aprof.items() if v[1]} if aprof else None,
        # the dominant kernel over the timed frames (HIP events)
        "timed_kernels": timed_kernels,
        # algorithmic bytes per launch of every instrumented kernel (attribution frames)
        "kernels_bytes_per_launch": {k: round(v[2] / v[1]) for k, v in aprof.items() if v[1] and v[2]}
        if aprof else None,
        "attribution_frames": f"{a.warmup}..{first - 1}" if n_attrib else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
