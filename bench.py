"""Benchmark: particle-substeps/s of the CK-MPM substep on one MI355X.

Workload (default): the 8M-particle mixed-material scene - 5M weakly compressible fluid
particles in a dam-break column plus a 3M-particle Drucker-Prager sand cone, sparse
512^3 grid (128^3 blocks), per-cell counts log-uniform in [4, 16], dt = 2e-5, FP32.
A step is one substep (grid update, fused G2P2G for both materials, rebin) over all
particles; inputs are resident in HBM before timing starts.  The timed substeps replay
hipGraphs of two substeps, captured before timing; the per-kernel durations come from a
second pass of as many substeps launched with HIP events at the phase boundaries.

Prints one JSON line.  roofline: the G2P2G phase (the dominant kernels), algorithmic
bytes per substep = sum_m N_m (2 S_m + 8) + 3584 G (state read + write, old-slot read and
move-code write per particle; velocity halo read 1536 B and mass/momentum write 2048 B
per grid block), divided by its device time measured with hipEvents on the engine
stream.  substep_hbm: the whole-substep byte model sum_m N_m (2 S_m + 12) + 7168 G.
fwd_bwd_over_fwd and batched_over_sequential: the other two ratios of the headline metric,
timed in the same run (cantilever with checkpoints every 25 substeps; the 64-scene fused
batch against sequential launches).  cpu_baseline: on the same scene shrunk 4x per axis
(1/64 of the particles), the FP64 oracle with OpenMP on all available cores (value), on one
core, and a NumPy restatement of the reference's Python demo substep.
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STATE_BYTES = {0: 48, 1: 48, 2: 52, 3: 16}   # x + F | x + F + logJp | x + J


def host_cpu() -> dict:
    """CPU model and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(workload: str, seconds: float = 8.0):
    """Three CPU baselines on the same sample of the workload (about `seconds` each):
    B1 the FP64 oracle on one core, B2 the same oracle with OpenMP on all the cores it
    may use (the headline CPU figure), B3 the NumPy restatement of the reference's Python
    CK-MPM demo substep (its own fixed-corotated material on its 256^3 grid, on the same
    particle positions)."""
    import numpy as np

    from ckmpm_amd import scenes
    from oracle.pic_numpy import PicDemo
    from oracle.pyoracle import Oracle, lib

    if workload == "mixed_dam_sand":
        sc = scenes.mixed_dam_sand(n_fluid=5_000_000 // 64, n_sand=3_000_000 // 64, scale=4)
        sample = "mixed_dam_sand shrunk 4x per axis: 128^3 cells, %d particles" % sc.n_particles
    else:
        sc = scenes.SCENES[workload]()
        sample = f"{workload}, {sc.n_particles} particles"
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))

    def oracle_rate(nthreads):
        lib().orc_set_threads(nthreads)
        o = Oracle.from_scene(sc)
        o.setup()
        o.substep(0.0, sc.dt)
        n, t0 = 0, time.perf_counter()
        while True:
            o.substep(sc.dt, sc.dt)
            n += 1
            el = time.perf_counter() - t0
            if el > seconds:
                break
        return sc.n_particles * n / el, n, el

    r1, n1, e1 = oracle_rate(1)
    r2, n2, e2 = oracle_rate(threads)
    pic = PicDemo(np.concatenate([m.x for m in sc.models]))
    n3, t0 = 0, time.perf_counter()
    while True:
        pic.substep(1e-4)
        n3 += 1
        e3 = time.perf_counter() - t0
        if e3 > seconds:
            break
    r3 = pic.x.shape[0] * n3 / e3
    unit = "particle-substeps/s"
    return {"value": r2, "unit": unit, "cores": threads, "kind": "port",
            "sample": f"{sample}; {n2} substeps of the FP64 oracle (OpenMP, {threads} threads) in {e2:.1f} s",
            "host_cpu": host_cpu(),
            "variants": {
                "B1_oracle_1_thread": {"value": r1, "unit": unit, "cores": 1, "kind": "port",
                                       "sample": f"{sample}; {n1} substeps in {e1:.1f} s"},
                "B2_oracle_openmp": {"value": r2, "unit": unit, "cores": threads, "kind": "port",
                                     "sample": f"{sample}; {n2} substeps in {e2:.1f} s"},
                "B3_python_demo_numpy": {"value": r3, "unit": unit, "cores": 1, "kind": "port",
                                         "sample": f"NumPy restatement of python-src/python_demo.py:101-215 (CK-MPM, "
                                                   f"fixed-corotated E 5e3, 256^3 grid, dt 1e-4) on the same "
                                                   f"{pic.x.shape[0]} positions; {n3} substeps in {e3:.1f} s"}}}


def copy_bandwidth(torch, mib: int = 1024, reps: int = 5) -> float:
    """On-box device-to-device copy bandwidth in GB/s (read + write bytes of a plain
    copy of `mib` MiB, best of `reps`, timed with events on torch's stream, which runs
    the copy): the attainable HBM rate the roofline fraction can also be read against."""
    n = mib << 18
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b) * 1e-3)
    del src, dst
    return 2.0 * n * 4 / best / 1e9


KERNEL_SOURCES = ["kernels.hip", "kernels.h", "material.h", "svd.h", "stencil.h", "common.h"]


def kernel_source_sha16() -> str:
    """Hash of the G2P2G kernel sources: ties a traffic measurement to the code it read."""
    import hashlib

    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "ckmpm_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _measured_traffic(path: str, workload: str):
    """HBM bytes per G2P2G launch pair for this workload, as measured with rocprofv3
    FETCH_SIZE + WRITE_SIZE and scaled by a plain copy's reading (None if absent), and
    whether that measurement was taken on the current kernel sources."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(workload)
    if e is not None:
        e = dict(e)
        e["current"] = e.get("kernel_source_sha16") == kernel_source_sha16()
    return e


def bench_batch(args, torch, only=None):
    """64 independent scenes: fused hipGraph batch vs one scene after the other (and the
    other batch modes unless `only` names the ones to run)."""
    from ckmpm_amd import scenes
    from ckmpm_amd.engine import MPMBatch, engine_from_scene

    t = time.perf_counter()
    scs = scenes.batch_scenes()
    engines = [engine_from_scene(sc) for sc in scs]
    N = sum(sc.n_particles for sc in scs)
    print(f"[bench] {len(scs)} scenes, {N} particles, set up in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    res = {}
    # batched = fused launches: each kernel of a substep runs once for all scenes, its
    # workgroups split by scene, two substeps per hipGraph replay; one hipGraph per scene
    # on the scene's own stream, one fork/join graph over all scenes and plain interleaved
    # streams are reported beside it
    modes = (("batched", MPMBatch.FUSED), ("engine_graphs", MPMBatch.ENGINE_GRAPHS), ("single_graph", MPMBatch.GRAPH),
             ("streams", MPMBatch.STREAMS), ("sequential", MPMBatch.SEQUENTIAL))
    for name, mode in modes:
        if only is not None and name not in only:
            continue
        b = MPMBatch(engines)
        b.Run(args.warmup, mode)
        tuned = None
        if mode == MPMBatch.FUSED:
            # the batch's own tuning (per kind G2P2G workgroup size, launch order), untimed
            b.autotune(2)
            tuned = {"serial_g2p2g": b.serial_g2p2g,
                     "g2p2g_512_by_kind": {str(k): sum(1 for e, sc in zip(engines, scs) for m in range(len(sc.models))
                                                       if sc.models[m].material.kind == k and e.tuning(m) == 512)
                                           for k in sorted({m.material.kind for sc in scs for m in sc.models})}}
        b.Synchronize()
        torch.cuda.synchronize()
        c0 = [e.counters() for e in engines]
        t0 = time.perf_counter()
        b.Run(args.steps, mode)
        b.Synchronize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        c1 = [e.counters() for e in engines]
        G = sum(0.5 * (a.neighbor_blocks + c.neighbor_blocks) for a, c in zip(c0, c1))
        res[name] = {"value": N * args.steps / el, "ms_per_step": 1e3 * el / args.steps, "G": G,
                     "captures": b.captures}
        if tuned is not None:
            res[name]["tuning"] = tuned
        print(f"[bench] {name}: {res[name]['ms_per_step']:.3f} ms per batch substep", file=sys.stderr, flush=True)
        b.close()
    substep_bytes = sum(m.x.shape[0] * (2 * STATE_BYTES[m.material.kind] + 12) for sc in scs for m in sc.models)
    substep_bytes += 7168 * res["batched"]["G"]
    ach = substep_bytes / (res["batched"]["ms_per_step"] * 1e-3) / 1e9
    kinds = [sc.models[0].material.kind for sc in scs]
    return {
        "metric": "particle-substeps/s batched (64 scenes, one fused launch per kernel and substep, hipGraph-replayed) vs sequential single-scene",
        "value": res["batched"]["value"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["batched"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 scene generator)",
        "config": {"workload": "batch64", "scenes": len(scs), "particles": N,
                   "grids": {str(g): sum(1 for sc in scs if sc.dim[0] * 4 == g) for g in (64, 128, 256)},
                   "materials": {str(k): kinds.count(k) for k in range(4)}, "dt": scs[0].dt},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole batched substep (byte model sum N (2S + 12) + 7168 G over all scenes)",
                     "bytes_per_batch_substep": substep_bytes},
        "sequential": {"value": res["sequential"]["value"], "ms_per_step": res["sequential"]["ms_per_step"]},
        "modes_ms_per_step": {k: v["ms_per_step"] for k, v in res.items()},
        "batched_over_sequential": res["batched"]["value"] / res["sequential"]["value"],
        "graph_captures": res["batched"]["captures"],
        "tuning": res["batched"].get("tuning"),
    }


def bench_adjoint(args, torch):
    """Differentiation workload: the cantilever (fixed-corotated, per-particle E, v0),
    300 substeps, checkpoints every 25; forward alone vs forward + backward."""
    import numpy as np

    from ckmpm_amd import scenes
    from ckmpm_amd.autodiff import clear_engine_cache, simulate
    from ckmpm_amd.engine import engine_from_scene

    sc = scenes.cantilever()
    steps = sc.substeps
    m = sc.models[0]
    N = m.x.shape[0]
    logE = torch.tensor(np.log(m.mu_p.astype(np.float64) * 2 * 1.3), requires_grad=True)
    v0 = torch.zeros((N, 3), dtype=torch.float64, requires_grad=True)
    # forward alone: plain substeps on an engine without a tape; the median over three
    # fresh engines (one 300-substep run is ~23 ms, so a single host hiccup shows)
    fwd_samples = []
    for _ in range(3):
        eng = engine_from_scene(sc)
        torch.cuda.synchronize()
        gc.collect()
        gc.disable()
        try:
            t0 = time.perf_counter()
            for k in range(steps):
                eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
            eng.Synchronize()
            fwd_samples.append(time.perf_counter() - t0)
        finally:
            gc.enable()
        eng.close()
    t_fwd = float(np.median(fwd_samples))
    target = torch.tensor([0.35, 0.5, 0.5], dtype=torch.float64)

    def fwd_bwd(segment):
        logE.grad = None
        v0.grad = None
        # the host loop launches every substep: a pass of Python's cyclic garbage
        # collector inside it (tens of ms) would stall the GPU, so collect first and
        # keep the collector off during the call (as timeit does)
        gc.collect()
        gc.disable()
        try:
            t0 = time.perf_counter()
            x = simulate(logE, v0, sc, steps, segment)
            loss = ((x.mean(0) - target) ** 2).sum()
            t1 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
        finally:
            gc.enable()
        print(f"[bench] cantilever {N} particles, segment {segment}: fwd {t_fwd * 1e3 / steps:.3f} ms/substep, "
              f"recorded fwd {(t1 - t0) * 1e3 / steps:.3f}, backward {(t2 - t1) * 1e3 / steps:.3f}",
              file=sys.stderr, flush=True)
        return t2 - t0, loss

    # Each mode: a first call (builds and sets up the engine), then the timed call, which
    # reuses that engine rewound to its setup state (as every iteration of a gradient
    # loop after the first does).  Full tape (every substep recorded in HBM, no replay),
    # then checkpoints every 25.
    # The timed figure of each mode is the median of three calls after the first.
    t_tape_first, _ = fwd_bwd(None)
    tape_samples = [fwd_bwd(None)[0] for _ in range(3)]
    t_tape = float(np.median(tape_samples))
    g_tape = (logE.grad.clone(), v0.grad.clone())
    t_first, _ = fwd_bwd(25)
    ck = [fwd_bwd(25) for _ in range(3)]
    ck_samples = [t for t, _ in ck]
    t_all, loss = float(np.median(ck_samples)), ck[-1][1]
    tape_vs_ck = max(float((g_tape[0] - logE.grad).norm() / logE.grad.norm()),
                     float((g_tape[1] - v0.grad).norm() / v0.grad.norm()))
    # the kept engine holds a stream on a hardware queue of its own, which slows other
    # work's graph replays while it exists: release it before anything else is timed
    clear_engine_cache()
    return {
        "metric": "particle-substeps/s fwd+bwd (cantilever, per-particle log E and v0 gradients)",
        "value": N * steps / t_all,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": 1e3 * t_all / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 scene generator)",
        "config": {"workload": "cantilever", "particles": N, "grid": "128x128x128 cells", "dt": sc.dt,
                   "substeps": steps, "checkpoint_every": 25, "material": "fixed-corotated"},
        "fwd": {"value": N * steps / t_fwd, "ms_per_step": 1e3 * t_fwd / steps,
                "samples_ms_per_step": [1e3 * t / steps for t in fwd_samples]},
        "fwd_bwd_over_fwd": t_all / t_fwd,
        "samples_ms_per_step": [1e3 * t / steps for t in ck_samples],
        "timing": "median of three calls after a first call (forward: of three fresh engines)",
        "first_call": {"ms_per_step": 1e3 * t_first / steps, "fwd_bwd_over_fwd": t_first / t_fwd,
                       "note": "includes building and setting up the engine"},
        "full_tape": {"value": N * steps / t_tape, "ms_per_step": 1e3 * t_tape / steps,
                      "fwd_bwd_over_fwd": t_tape / t_fwd, "grad_rel_diff_vs_checkpointed": tape_vs_ck,
                      "samples_ms_per_step": [1e3 * t / steps for t in tape_samples],
                      "first_call_ms_per_step": 1e3 * t_tape_first / steps},
        "loss": float(loss.detach()),
        "grad_norms": {"logE": float(logE.grad.norm()), "v0": float(v0.grad.norm())},
    }


def bench_hash(args, torch):
    """Sparse block table: build from the exterior block list of the 8M mixed scene
    (after warm-up substeps) and Find throughput on device keys, against a dense gather."""
    import numpy as np

    from ckmpm_amd import scenes
    from ckmpm_amd.engine import engine_from_scene
    from ckmpm_amd.hashtable import BlockHashTable

    sc = scenes.mixed_dam_sand()
    eng = engine_from_scene(sc)
    for k in range(4):
        eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
    keys, (A, Nb, Ex) = eng.layout(0)
    keys = keys[:Ex].astype(np.uint32)
    eng.close()
    nblocks = int(np.prod(sc.dim))
    h = BlockHashTable(sc.dim, Ex)
    dk = torch.from_numpy(keys.view(np.int32)).cuda()
    dv = torch.arange(Ex, dtype=torch.int32, device="cuda")
    h.InsertDevice(dk.data_ptr(), dv.data_ptr(), Ex)   # warm-up
    reps = 20
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        h.InsertDevice(dk.data_ptr(), dv.data_ptr(), Ex)
    t_build = (time.perf_counter() - t0) / reps
    inf = h.info()
    # queries: the 27-neighbourhood lookups of the rebin, repeated to 8M keys
    rng = np.random.default_rng(1)
    q = keys[rng.integers(0, Ex, 8 << 20)]
    dq = torch.from_numpy(q.view(np.int32)).cuda()
    out = torch.empty_like(dq)
    dense = torch.full((nblocks,), -1, dtype=torch.int32, device="cuda")
    dense[dk.long()] = dv
    res = {}
    forms = {"cooperative": h.COOPERATIVE, "scalar": h.SCALAR, "lds": h.LDS}
    names = ["cooperative", "scalar", "dense"] + (["lds"] if h.info().buckets <= 2560 else [])
    for name in names:
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                if name == "dense":
                    out = dense[dq.long()]
                else:
                    h.FindDevice(dq.data_ptr(), dq.numel(), out.data_ptr(), cooperative=forms[name])
            h.Synchronize()
            torch.cuda.synchronize()
            el = (time.perf_counter() - t0) / reps
        res[name] = dq.numel() / el
    ok = bool((out.cpu().numpy() >= 0).all())
    h.close()
    return {
        "metric": "block-table lookups/s (sparse hash vs dense table) on the 8M mixed scene's block list",
        "value": res["cooperative"],
        "unit": "lookups/s",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": 1e3 * dq.numel() / res["cooperative"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded splitmix64 scene generator)",
        "config": {"workload": "hash", "keys": int(Ex), "domain_blocks": nblocks, "queries": int(dq.numel())},
        "lookups_per_s": res,
        "build_ms": 1e3 * t_build,
        "occupancy": {"buckets": inf.buckets, "occupied_slots": inf.occupied_slots, "stash": inf.stash,
                      "fill_histogram": list(inf.fill_histogram)},
        "dense_table_bytes": 4 * nblocks, "hash_bytes": 8 * 16 * inf.buckets,
        "all_found": ok,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="mixed_dam_sand")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the fwd+bwd and batched ratios that the default run also times")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB three times (HBM counter calibration)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="measured HBM bytes per G2P2G launch (from rocprofv3 FETCH_SIZE / WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); the multi-GPU sharding of the "
              "substep is described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ckmpm_amd import scenes
    from ckmpm_amd.engine import MPMBatch, engine_from_scene

    if args.workload == "batch64":
        print(json.dumps(bench_batch(args, torch)), flush=True)
        return
    if args.workload == "hash":
        print(json.dumps(bench_hash(args, torch)), flush=True)
        return
    if args.workload == "cantilever":
        print(json.dumps(bench_adjoint(args, torch)), flush=True)
        return

    if args.copy_calibration_mib > 0:
        src = torch.ones(args.copy_calibration_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: 3 x {args.copy_calibration_mib} MiB", file=sys.stderr, flush=True)
        del src, dst

    t = time.perf_counter()
    sc = scenes.SCENES[args.workload]()
    print(f"[bench] scene {sc.name}: {sc.n_particles} particles, {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    eng = engine_from_scene(sc)
    print(f"[bench] engine set up, {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    dt = sc.dt
    # warm-up: the autotuner's substeps (two candidate workgroup sizes) count toward it
    tune = max(1, args.warmup // 2 - 2)
    eng.autotune(tune)
    for k in range(max(0, args.warmup - 2 * (tune + 1))):
        eng.Substep(dt, dt)
    eng.Synchronize()
    torch.cuda.synchronize()
    print(f"[bench] {args.warmup} warmup substeps done, timing {args.steps}", file=sys.stderr, flush=True)
    # the job: K substeps replayed as hipGraphs of two substeps (one launch per kernel),
    # captured before the timed region; the K substeps are one call into the library
    b = MPMBatch([eng])
    b.Run(2, MPMBatch.FUSED)
    b.Synchronize()
    c0 = eng.counters()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.Run(args.steps, MPMBatch.FUSED)
    b.Synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    c1 = eng.counters()
    b.close()
    elapsed = t1 - t0
    # per-kernel durations: K more substeps as plain launches on the engine stream with
    # HIP events at the phase boundaries (the events add gaps between launches, so this
    # pass is not the throughput)
    eng.set_profiling(True)
    for _ in range(args.steps):
        eng.Substep(dt, dt)
    eng.Synchronize()
    (ms_grid, ms_g2p2g, ms_rebin), nsub = eng.phase_times()
    c2 = eng.counters()
    eng.set_profiling(False)
    N = sc.n_particles
    G = 0.5 * (c0.neighbor_blocks + c1.neighbor_blocks)
    G_ev = 0.5 * (c1.neighbor_blocks + c2.neighbor_blocks)
    g2p2g_bytes = sum(m.x.shape[0] * (2 * STATE_BYTES[m.material.kind] + 8) for m in sc.models) + 3584 * G_ev
    g2p2g_s = ms_g2p2g / max(nsub, 1) * 1e-3
    achieved = g2p2g_bytes / g2p2g_s / 1e9
    substep_bytes = sum(m.x.shape[0] * (2 * STATE_BYTES[m.material.kind] + 12) for m in sc.models) + 7168 * G
    sps = args.steps / elapsed
    out = {
        "metric": ("particle-substeps/s fwd at 8M particles (mixed fluid + sand, 512^3)"
                   if args.workload == "mixed_dam_sand" else
                   f"particle-substeps/s fwd ({sc.name}, {N} particles, "
                   f"{4 * sc.dim[0]}x{4 * sc.dim[1]}x{4 * sc.dim[2]} cells)"),
        "value": N * args.steps / elapsed,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 scene generator)",
        "config": {"workload": sc.name, "particles": N, "grid": "%dx%dx%d cells" % tuple(4 * d for d in sc.dim),
                   "dt": dt, "models": [m.material.kind for m in sc.models], "substeps_timed": args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "G2P2GKernel (fluid + Drucker-Prager launches per substep)",
                     "bytes_per_launch_pair": g2p2g_bytes, "avg_ms": g2p2g_s * 1e3},
        "substep_hbm": {"bytes": substep_bytes, "frac": substep_bytes * sps / (HBM_PEAK_GBS * 1e9)},
        "phase_ms_per_substep": {"grid_update": ms_grid / max(nsub, 1), "g2p2g": ms_g2p2g / max(nsub, 1),
                                 "rebin": ms_rebin / max(nsub, 1)},
        "timing": {"value": f"{args.steps} substeps replayed as two-substep hipGraphs (fused launches)",
                   "phases": f"{nsub} further substeps as plain launches with HIP events at phase boundaries",
                   "events_ms_per_step": (ms_grid + ms_g2p2g + ms_rebin) / max(nsub, 1)},
        "blocks": {"active": c1.active_blocks, "neighbor": c1.neighbor_blocks, "exterior": c1.exterior_blocks},
        "g2p2g_workgroup": [eng.tuning(m) for m in range(len(sc.models))],
    }
    traffic = _measured_traffic(args.traffic, sc.name)
    if traffic is not None:
        # a measurement of other kernel sources is reported beside, not as the traffic
        out["roofline"]["traffic"] = traffic["bytes_per_launch_pair"] if traffic["current"] else None
        out["roofline"]["traffic_source"] = traffic["source"]
        out["roofline"]["traffic_kernel_source_sha16"] = traffic.get("kernel_source_sha16")
        if not traffic["current"]:
            out["roofline"]["traffic_measured_on_other_sources"] = traffic["bytes_per_launch_pair"]
    out["roofline"]["kernel_source_sha16"] = kernel_source_sha16()
    eng.close()
    copy_gbs = copy_bandwidth(torch)
    torch.cuda.empty_cache()  # the copy calibration's buffers: nothing else holds HBM now
    out["roofline"]["copy_gbs"] = copy_gbs
    out["roofline"]["frac_of_copy"] = achieved / copy_gbs
    if args.workload == "mixed_dam_sand" and not args.no_extras:
        # the other two ratios of the headline metric, timed in the same run: forward +
        # backward against forward on the differentiation workload (checkpoints every
        # 25 substeps) and the fused 64-scene batch against sequential single-scene launches
        ad = bench_adjoint(args, torch)
        out["fwd_bwd_over_fwd"] = ad["fwd_bwd_over_fwd"]
        out["fwd_bwd"] = {"workload": "cantilever", "particles": ad["config"]["particles"],
                          "fwd_ms_per_step": ad["fwd"]["ms_per_step"], "fwd_bwd_ms_per_step": ad["ms_per_step"],
                          "fwd_samples_ms_per_step": ad["fwd"]["samples_ms_per_step"],
                          "fwd_bwd_samples_ms_per_step": ad["samples_ms_per_step"],
                          "checkpoint_every": 25, "timing": ad["timing"]}
        bargs = argparse.Namespace(**{**vars(args), "steps": 40, "warmup": 4})
        bt = bench_batch(bargs, torch, only=("batched", "sequential"))
        out["batched_over_sequential"] = bt["batched_over_sequential"]
        out["batched"] = {"workload": "batch64", "scenes": bt["config"]["scenes"], "particles": bt["config"]["particles"],
                          "batched_ms_per_step": bt["ms_per_step"], "sequential_ms_per_step": bt["sequential"]["ms_per_step"],
                          "value": bt["value"], "substeps": 40, "tuning": bt["tuning"]}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
