"""Benchmark: gravity on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 2] [--no-cpu] [--only LEG]

Prints ONE JSON line.  The headline (value, roofline, cpu_baseline) is the direct sum at
N = 262,144 (two Plummer spheres): a step is one fused force + kick-drift pass over all N
bodies (N^2 pairwise interactions, 20 flops each by the usual convention).  Inputs are
resident in HBM before the timed region; the region is bracketed by device
synchronisation, and the force kernel's own duration is measured with HIP events on the
library's stream.  The CPU baseline (oracle restatement of the reference brute-force
loop) is timed on a bounded subset of rows on this host's cores.

Secondary objects in the same line:
  "c1"         the reference's CPU-runnable case: N = 4096 Plummer, 100 steps, GPU and the
               CPU brute-force restatement over all 100 steps;
  "batched"    512 Plummer systems, N log-uniform in [256, 32768], 50 steps replayed from
               a hipGraph, against the same systems stepped one launch per system;
  "barnes_hut" N = 2^20 exponential disk + Plummer bulge, theta 0.5, leaf 16, 10 steps,
               with the reference Barnes-Hut restatement timed on the CPU;
  "bh_batched" the served workload's Barnes-Hut jobs in one Barnes-Hut batch (one sort,
               build and walk per step for all of them), against the same systems one
               context at a time;
  "served"     a Poisson stream of jobs through the continuous-batching server: served
               steps/s and p50/p99 latency.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G-interactions/s + % FP32 peak at N=262144; served steps/s, p50/p99 latency"
FP32_PEAK_TFLOPS = 157.3          # MI355X published FP32 vector peak
FLOPS_PER_INTERACTION = 20        # GPU Gems 3 ch. 31 convention
N_C2 = 262144
EPS_C2 = 0.005
DT_C2 = 5e-4
SEED = 2026
CPU_THREADS_MAX = 16              # the host CPU share of one GPU on the benchmark hosts
CPU_TARGET_S = 10.0
HBM_PEAK_GBS = 8000.0              # MI355X published HBM3E bandwidth


def cpu_threads() -> int:
    """Threads for the CPU baselines: every core this process may run on, capped at 16,
    the CPU share of one GPU on the benchmark hosts (the host's other cores belong to
    the other GPUs' jobs)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(CPU_THREADS_MAX, avail))


def cpu_facts() -> dict:
    """What the CPU baseline ran on: model, logical CPUs, usable CPUs, threads used."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": avail, "threads": cpu_threads(),
            "thread_cap": f"{CPU_THREADS_MAX} (one GPU's CPU share on the benchmark hosts)"}


def cpu_baseline(bodies: np.ndarray, eps: float) -> dict:
    """Oracle restatement of the reference brute-force loop (source/solvers/
    cpu_brute_force.h:26-40) on a seeded subset of rows x all N columns."""
    from oracle import oracle as O  # the CPU baseline leg is the only oracle user here
    if not (os.path.exists(O.LIB_PATH) and os.path.exists(O.LIB_O3_PATH)):
        O.build(ref=False)
    threads = cpu_threads()
    n = bodies.shape[0]
    rng = np.random.default_rng(SEED)
    work = bodies.copy()
    rows = 256
    t = 0.0
    while True:
        sel = rng.choice(n, rows, replace=False).astype(np.uint32)
        t0 = time.perf_counter()
        O.brute_force_rows(work, sel, G=1.0, eps=eps, threads=threads)
        t = time.perf_counter() - t0
        if t >= 0.5 or rows >= n:
            break
        rows = min(n, rows * 4)
    # scale the final sample to roughly CPU_TARGET_S
    rows = int(min(n, max(rows, rows * CPU_TARGET_S / max(t, 1e-6))))
    sel = rng.choice(n, rows, replace=False).astype(np.uint32)
    t0 = time.perf_counter()
    O.brute_force_rows(work, sel, G=1.0, eps=eps, threads=threads)
    t = time.perf_counter() - t0
    rate = rows * n / t
    # the same rows with the -O3 -march=x86-64-v3 build
    O.variant("o3")
    try:
        t0 = time.perf_counter()
        O.brute_force_rows(work, sel, G=1.0, eps=eps, threads=threads)
        t3 = time.perf_counter() - t0
    finally:
        O.variant("o2")
    return {"value": rate / 1e9, "unit": "G-interactions/s", "cores": threads, "kind": "port",
            "sample": f"{rows} seeded rows x all {n} bodies, 1 force pass, {t:.1f} s, "
                      f"oracle/nbody_oracle.c brute force (-O2, {threads} threads)",
            "o3_march_x86_64_v3": {"value": rows * n / t3 / 1e9, "seconds": t3}, **cpu_facts()}


def load_traffic(name: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def momentum_drift(d0, d1) -> float:
    """|P_end - P_start| over the momentum scale sqrt(2 M KE_start) (= M v_rms), from two
    nbhip_diagnostics results (KE, PE, P xyz, L xyz, M)."""
    dp = np.linalg.norm(np.asarray(d1[2:5], np.float64) - np.asarray(d0[2:5], np.float64))
    return float(dp / np.sqrt(2.0 * d0[8] * d0[0]))


def run_direct(args) -> dict:
    from nbody_amd import _lib, ics
    n = N_C2
    bodies = ics.two_plummer(n, seed=SEED)
    p = _lib.params(dt=DT_C2, eps=EPS_C2, size=10000.0, wrap=1, G=1.0)
    ctx = _lib.Context(0)
    ctx.upload_bodies(bodies)
    _, tuned = ctx.tune(p)          # the autotuner picks the launch shape for this N
    e0 = ctx.diagnostics(p)
    for _ in range(args.warmup):
        ctx.step(p, 1)
    ctx.synchronize()

    t0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.step(p, 1)
    ms_events, launches = ctx.timer_stop()
    ctx.synchronize()
    wall = time.perf_counter() - t0

    e1 = ctx.diagnostics(p)
    # accuracy: forces at the final positions vs FP64 on 1024 seeded bodies
    ctx.accelerate(p)
    sample = np.random.default_rng(SEED).choice(n, 1024, replace=False).astype(np.uint32)
    _, rel = ctx.force_error_f64(p, sample)
    ctx.close()

    inter = float(n) * n
    per_launch_ms = ms_events / max(launches, 1)
    value = inter * args.steps / (ms_events / 1e3) / 1e9
    achieved_tflops = FLOPS_PER_INTERACTION * inter / (per_launch_ms / 1e3) / 1e12
    E0 = e0[0] + e0[1]
    E1 = e1[0] + e1[1]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "G-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_events / args.steps,
        "wall_ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "pct_fp32_peak": 100.0 * achieved_tflops / FP32_PEAK_TFLOPS,
        "config": {
            "workload": "direct-sum two-Plummer collision, N=262144, eps=0.005, dt=5e-4, "
                        "fused force+kick-drift step, FP32 with TwoSum-compensated force sum",
            "n_bodies": n, "eps": EPS_C2, "dt": DT_C2, "seed": SEED,
            "launch": {"P": tuned[0], "W": tuned[1], "jsplit": tuned[2], "pair_symmetric": tuned[2] == 0,
                       "autotuned": True},
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS,
            "traffic": load_traffic("traffic_direct_sym_n262144.json" if tuned[2] == 0
                                    else "traffic_direct_n262144.json"),
            "kernel": ("k_force_sym + k_reduce_slots (one fused step: each pair once, then each "
                       "body's per-block partial sums added and the kick-drift)" if tuned[2] == 0
                       else "k_force (fused step)"),
            "kernel_ms": per_launch_ms,
            "launches": launches,
        },
        "accuracy": {
            "rel_err_vs_f64_max": float(rel.max()),
            "rel_err_vs_f64_median": float(np.median(rel)),
            "sample": "1024 seeded bodies, FP64 direct sum on the device",
            "energy_drift": float((E1 - E0) / abs(E0)),
            "momentum_drift": momentum_drift(e0, e1),
            "energy_steps": args.warmup + args.steps,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(bodies, EPS_C2)
    return out


def run_c1(args) -> dict:
    """The reference's CPU-runnable case: N = 4096 Plummer (a = 1, M = 1, G = 1), eps 0.01,
    dt 1e-3, wrap off, 100 steps, on the GPU (autotuned launch shape; steps enqueued one
    call per step, and replayed from a captured hipGraph) and on the CPU (the brute-force
    restatement and the reference's integrator over all 100 steps).  At this size the
    force pass is launch- and occupancy-bound: 32 blocks of 128 bodies per j-chunk."""
    from nbody_amd import _lib, ics
    n, steps = 4096, 100
    bodies = ics.plummer(n, seed=SEED)
    p = _lib.params(dt=1e-3, eps=0.01, size=10000.0, wrap=0, G=1.0)
    ctx = _lib.Context(0)
    ctx.upload_bodies(bodies)
    _, tuned = ctx.tune(p)
    ctx.upload_bodies(bodies)
    ctx.step(p, 2)
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        ctx.step(p, 1)
    ms, launches = ctx.timer_stop()
    ctx.upload_bodies(bodies)
    ctx.step_graph(p, steps)          # capture (untimed)
    ctx.upload_bodies(bodies)
    ctx.synchronize()
    ctx.timer_start()
    ctx.step_graph(p, steps)
    ms_graph, _ = ctx.timer_stop()
    gpu = ctx.download_bodies()
    ctx.close()
    inter = float(n) * n * steps
    out = {
        "workload": "N=4096 Plummer (a=1, M=1, G=1), eps=0.01, dt=1e-3, wrap off, 100 steps",
        "value": inter / (ms / 1e3) / 1e9, "unit": "G-interactions/s",
        "ms_per_step": ms / steps, "steps_per_s": steps / (ms / 1e3),
        "graph": {"value": inter / (ms_graph / 1e3) / 1e9, "ms_per_step": ms_graph / steps},
        "pct_fp32_peak": 100.0 * FLOPS_PER_INTERACTION * inter / (ms / 1e3) / (FP32_PEAK_TFLOPS * 1e12),
        "launch": {"P": tuned[0], "W": tuned[1], "jsplit": tuned[2], "pair_symmetric": tuned[2] == 0,
                   "autotuned": True},
        "bound": "launch/occupancy: 4096 bodies give 32 i-blocks of 128",
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        cpu = bodies.copy()
        t0 = time.perf_counter()
        for _ in range(steps):
            O.brute_force(cpu, eps=0.01, threads=threads)
            O.integrate(cpu, np.float32(1e-3), 10000.0, False, threads=threads)
        t = time.perf_counter() - t0
        dx = float(np.abs(gpu[:, 0:3].astype(np.float64) - cpu[:, 0:3]).max())
        out["cpu_baseline"] = {"value": inter / t / 1e9, "unit": "G-interactions/s", "cores": threads,
                               "kind": "port", "seconds": t, "steps_per_s": steps / t,
                               "sample": "all 100 steps: oracle/nbody_oracle.c brute force + integrate (-O2)",
                               **cpu_facts()}
        out["max_abs_dx_gpu_vs_cpu_after_100_steps"] = dx
    return out


def run_integrate(args) -> dict:
    """The standalone semi-implicit Euler kick-drift (nbhip_integrate, k_integrate) on
    N = 2^24 device-drawn Plummer bodies, with and without the torus wrap: 80 B of HBM
    traffic per body (read position, velocity, acceleration; write position, velocity)
    against the 8 TB/s peak.  Inside a step it is fused into the force pass; this is the
    Solver::integrate path on its own."""
    from nbody_amd import _lib
    n = 1 << 24
    reps = 20
    ctx = _lib.Context(0)
    ctx.generate(_lib.ics(_lib.IC_PLUMMER, seed=SEED), n)
    out = {"workload": f"k_integrate alone, N = 2^24 Plummer bodies drawn on the device, {reps} launches",
           "bodies": n, "bytes_per_body": 80}
    for name, wrap in (("wrap", 1), ("no_wrap", 0)):
        p = _lib.params(dt=1e-6, eps=0.01, size=10000.0, wrap=wrap)
        ctx.integrate(p)  # warm
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(reps):
            ctx.integrate(p)
        ms, _ = ctx.timer_stop()
        per = ms / reps
        gbs = 80.0 * n / (per * 1e-3) / 1e9
        out[name] = {"ms_per_launch": per, "GB/s": gbs, "frac_of_hbm_peak": gbs / 8000.0}
    ctx.close()
    return out


def run_batch(args) -> dict:
    """512 independent Plummer systems (N log-uniform in [256, 32768], own seeds), eps 0.01,
    dt 1e-3: autotuned batched launch, 50 steps captured in a hipGraph and replayed."""
    from nbody_amd import _lib, ics
    rng = np.random.default_rng(SEED)
    ns = np.exp(rng.uniform(np.log(256), np.log(32768), 512)).astype(int)
    p = _lib.params(dt=1e-3, eps=0.01, size=10000.0, wrap=1)
    ctx = _lib.Context(0)
    batch = _lib.Batch(ctx)
    systems = []
    for k, n in enumerate(ns):
        b = ics.plummer(int(n), seed=SEED + k)
        systems.append(b)
        batch.add(b, p)
    tuned_ms = batch.tune()
    info = batch.info()
    steps = 50
    batch.step(steps, _lib.BATCH_GRAPH)          # capture + first replay (untimed)
    ctx.synchronize()
    ctx.timer_start()
    batch.step(steps, _lib.BATCH_GRAPH)
    ms, _ = ctx.timer_stop()
    inter = info.interactions_per_step
    rate = inter * steps / (ms / 1e3)
    # the unbatched path: the same systems, one launch (+ reduce) per system per step, with
    # the single-system launch shape (j-split to fill the GPU); and "plain", one launch per
    # system with no j-split
    base_steps = 5
    batch.step(1, _lib.BATCH_PER_SYSTEM)
    ctx.synchronize()
    ctx.timer_start()
    batch.step(base_steps, _lib.BATCH_PER_SYSTEM)
    ms_base, _ = ctx.timer_stop()
    rate_base = inter * base_steps / (ms_base / 1e3)
    os.environ["NBHIP_DIRECT_JSPLIT"] = "1"
    try:
        ctx.timer_start()
        batch.step(base_steps, _lib.BATCH_PER_SYSTEM)
        ms_plain, _ = ctx.timer_stop()
    finally:
        del os.environ["NBHIP_DIRECT_JSPLIT"]
    rate_plain = inter * base_steps / (ms_plain / 1e3)
    batch.close()
    ctx.close()
    out = {
        "workload": "512 Plummer systems, N log-uniform in [256, 32768], seeds 2026+k, "
                    "eps=0.01, dt=1e-3, 50 steps in one hipGraph",
        "systems": int(info.systems), "bodies": int(info.bodies),
        "interactions_per_step": inter,
        "value": rate / 1e9, "unit": "G-interactions/s",
        "ms_per_step": ms / steps,
        "pct_fp32_peak": 100.0 * FLOPS_PER_INTERACTION * rate / (FP32_PEAK_TFLOPS * 1e12),
        "launch": {"P": info.P, "W": info.W,
                   "mode": {0: "grid", 1: "persistent", 2: "pairs"}.get(info.launch, info.launch),
                   "autotuned_ms_per_step": tuned_ms},
        "unbatched_value": rate_base / 1e9,
        "unbatched_ms_per_step": ms_base / base_steps,
        "speedup_vs_unbatched": rate / rate_base,
        "unbatched_plain_value": rate_plain / 1e9,
        "speedup_vs_unbatched_plain": rate / rate_plain,
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        pick = np.random.default_rng(SEED).choice(len(systems), 64, replace=False)
        def timed(variant: str) -> tuple[float, float]:
            O.variant(variant)
            try:
                t0 = time.perf_counter()
                work = 0.0
                for k in pick:
                    b = systems[k].copy()
                    O.brute_force(b, eps=0.01, threads=threads)
                    work += float(b.shape[0]) ** 2
                return work, time.perf_counter() - t0
            finally:
                O.variant("o2")
        work, t = timed("o2")
        _, t3 = timed("o3")
        out["cpu_baseline"] = {"value": work / t / 1e9, "unit": "G-interactions/s", "cores": threads,
                               "kind": "port",
                               "sample": f"one full force pass of 64 of the 512 systems (seeded pick), "
                                         f"{t:.1f} s, -O2",
                               "o3_march_x86_64_v3": {"value": work / t3 / 1e9, "seconds": t3},
                               **cpu_facts()}
    return out


def run_bh(args) -> dict:
    """N = 2^20 (--bh-bodies) exponential disk (R_d 3, z_0 0.3, 80% of the mass) + Plummer
    bulge (a 0.5), theta 0.5 (reference criterion), leaf 16, eps 0.01, dt 1e-3, 10 steps,
    root cube 128."""
    from nbody_amd import _lib, ics
    n = max(2, args.bh_bodies)
    size = 128.0
    bodies = ics.disk_bulge(n, seed=SEED)
    p = _lib.params(mode=_lib.MODE_NLOGN, theta=0.5, leaf_size=16, eps=0.01, dt=1e-3, size=size,
                    wrap=1)
    ctx = _lib.Context(0)
    ctx.upload_bodies(bodies)
    e0 = ctx.diagnostics(p)
    # the textbook criterion at the same theta first, for comparison; then the reference's
    pt = _lib.params(mode=_lib.MODE_NLOGN, theta=0.5, leaf_size=16, eps=0.01, dt=1e-3, size=size,
                     wrap=1, opening=_lib.OPEN_TEXTBOOK)
    sample = np.random.default_rng(SEED).choice(n, min(n, 1024), replace=False).astype(np.uint32)
    ctx.accelerate(pt)
    _, rel_text = ctx.force_error_f64(pt, sample)
    _, inter = ctx.bh_interactions(p)    # also leaves the first pass's accelerations
    _, rel = ctx.force_error_f64(p, sample)
    steps = 10
    ctx.step(p, 1)
    ctx.synchronize()
    ctx.timer_start()
    ctx.step(p, steps)                   # eager: ~20 launches per step
    ms, _ = ctx.timer_stop()
    ctx.step_graph(p, steps)             # capture (untimed) ...
    ctx.synchronize()
    ctx.timer_start()
    ctx.step_graph(p, steps)             # ... and replay
    ms_graph, _ = ctx.timer_stop()
    e1 = ctx.diagnostics(p)
    # the phases of a step, timed by HIP events on the context's stream
    build_ms, walk_ms = ctx.bh_profile(p, steps)
    passes = ctx.bh_sort_passes()        # radix passes the profiled steps' sorts ran
    import ctypes
    cnt = ctypes.c_size_t()
    ctx._check(ctx.L.nbhip_bh_nodes(ctx.h, ctypes.byref(p), None, 0, ctypes.byref(cnt)))
    nodes = int(cnt.value)
    ctx.close()
    # algorithmic bytes of one build per body: keys 16 in + 12 out; radix sort 8 (the
    # digit counts of all passes in one read) + 24 per pass run (8 for a full sort; the
    # passes of skipped low digits read nothing), then the run fix-up's read of the keys (8);
    # levels 8 in + ~1.07 out; gather (position, velocity, caller's index and the sort's
    # value) 40 in + 36 out; moment scan 16 in + 32 out; next-smaller 4 out; node count and
    # emit ~5 (per-position counts and kept flags) -- 186 B + 24 per pass -- plus the 32 B
    # of every node written
    per_body = 186.0 + 24.0 * passes
    build_bytes = per_body * n + 32.0 * nodes
    build_gbs = build_bytes / (build_ms * 1e-3) / 1e9
    walk_tflops = FLOPS_PER_INTERACTION * inter / (walk_ms * 1e-3) / 1e12
    out = {
        "workload": f"Barnes-Hut N={n} disk+bulge, theta=0.5 (reference criterion), leaf 16, "
                    "eps=0.01, dt=1e-3, 10 steps, tree rebuilt every step",
        "steps_per_s": steps / (ms / 1e3), "ms_per_step": ms / steps,
        "ms_per_step_graph": ms_graph / steps,
        "interactions_per_step": int(inter),
        "value": inter * steps / (ms / 1e3) / 1e9, "unit": "G-interactions/s",
        "roofline": {
            "build": {"bound": "hbm", "achieved": build_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": build_gbs / HBM_PEAK_GBS, "ms": build_ms,
                      "algorithmic_bytes": build_bytes, "nodes": nodes,
                      "radix_passes": passes,
                      "bytes_formula": f"{per_body:.0f} B per body (186 + 24 per radix pass run) + 32 B per node"},
            "walk": {"bound": "valu", "achieved": walk_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": walk_tflops / FP32_PEAK_TFLOPS, "ms": walk_ms,
                     "flops": f"{FLOPS_PER_INTERACTION} per body-node interaction"},
            "timing": f"HIP events around each phase of {steps} fused steps (nbhip_bh_profile)",
        },
        "energy_drift": float(((e1[0] + e1[1]) - (e0[0] + e0[1])) / abs(e0[0] + e0[1])),
        "accuracy": {
            "rel_err_vs_f64_median": float(np.median(rel)),
            "rel_err_vs_f64_p99": float(np.quantile(rel, 0.99)),
            "rel_err_vs_f64_max": float(rel.max()),
            "sample": "1024 seeded bodies at the initial positions, FP64 direct sum on the device: "
                      "the approximation error of the reference's criterion (a cell is taken whole "
                      "when |com - x| > theta * edge) at theta 0.5, not rounding",
            "textbook_opening": {"rel_err_vs_f64_median": float(np.median(rel_text)),
                                 "rel_err_vs_f64_p99": float(np.quantile(rel_text, 0.99)),
                                 "rel_err_vs_f64_max": float(rel_text.max()),
                                 "criterion": "|com - x| > edge / theta, theta 0.5"},
            "energy_drift": float(((e1[0] + e1[1]) - (e0[0] + e0[1])) / abs(e0[0] + e0[1])),
            "momentum_drift": momentum_drift(e0, e1),
            "energy_steps": 1 + 3 * steps,
        },
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        def timed(variant: str):
            O.variant(variant)
            try:
                b = bodies.copy()
                t0 = time.perf_counter()
                nodes = O.bh_build(b, size)
                t_build = time.perf_counter() - t0
                t0 = time.perf_counter()
                _, v = O.bh_accelerate(b, size, theta=0.5, eps=0.01, threads=threads, nodes=nodes,
                                       visits=True)
                return t_build, time.perf_counter() - t0, v
            finally:
                O.variant("o2")
        t_build, t_walk, v = timed("o2")
        t3_build, t3_walk, _ = timed("o3")
        out["cpu_baseline"] = {"value": 1.0 / (t_build + t_walk), "unit": "steps/s", "cores": threads,
                               "kind": "port", "build_ms": t_build * 1e3, "walk_ms": t_walk * 1e3,
                               "interactions_per_step": int(v.sum()),
                               "sample": "1 force pass of the reference octree restatement "
                                         "(serial build, leaf 1, threaded walk), -O2",
                               "o3_march_x86_64_v3": {"value": 1.0 / (t3_build + t3_walk),
                                                      "build_ms": t3_build * 1e3,
                                                      "walk_ms": t3_walk * 1e3},
                               **cpu_facts()}
    return out


def run_served(args) -> dict:
    """Poisson stream of simulation jobs (Plummer or disk+bulge, N log-uniform in [1k, 256k],
    steps uniform in [10, 1000]; direct sum below 16k bodies, Barnes-Hut theta 0.5 above),
    device-resident, admitted against a 200 GB budget, continuously batched; each job's
    bodies are drawn on the device at admission (nbhip_generate), so none crosses PCIe on
    the way in.  Capacity is measured first with every job queued at t=0; the open stream
    then runs at 0.7x it."""
    from nbody_amd import serve
    count = 160
    jobs = serve.make_jobs(count, seed=SEED, device_ics=True)
    with serve.Server(hbm_budget=200e9) as srv:
        sat = srv.run(jobs)
        capacity = sat.jobs / sat.wall_s
        rate = 0.7 * capacity
        serve.poisson_arrivals(jobs, rate=rate, seed=SEED)
        rep = srv.run(jobs)
        leaves = {f"2^{c}": leaf for c, leaf in sorted(srv.leaf_by_class.items())}
    # the same arrivals under one first-come-first-served queue (the earlier policy)
    with serve.Server(hbm_budget=200e9, admission="fcfs") as srv:
        fcfs = srv.run(jobs).summary()
    # the Barnes-Hut jobs each on a context of a pool of 8 instead of in the Barnes-Hut
    # batch: saturated, then the same arrivals
    arrivals = [j.arrival for j in jobs]
    with serve.Server(hbm_budget=200e9, bh_mode="contexts") as srv:
        for j in jobs:
            j.arrival = 0.0
        ctx_sat = srv.run(jobs)
        for j, a in zip(jobs, arrivals):
            j.arrival = a
        ctx_rep = srv.run(jobs).summary()
    # the same jobs one at a time through the plain single-system path; the Barnes-Hut jobs
    # take the leaf sizes the server's autotuner picked
    for j in jobs:
        if j.mode == 1:
            j.params.leaf_size = leaves.get(f"2^{max(0, int(j.n).bit_length() - 1)}", j.params.leaf_size)
    seq = serve.run_sequential(jobs)
    s = rep.summary()
    n_direct = sum(1 for j in jobs if j.mode == 0)
    return {
        "workload": f"{count} jobs: Plummer or disk+bulge, N log-uniform in [1024, 262144], steps "
                    "uniform in [10, 1000]; direct sum below 16384 bodies (batched), Barnes-Hut "
                    "theta 0.5 above (batched: one sort, build and walk per step for all of them, "
                    "on a stream of its own), leaf size autotuned per power-of-two size class, "
                    "chunks of 16 steps replayed from captured graphs; bodies drawn on the device "
                    "at admission; "
                    "200 GB budget; "
                    f"Poisson arrivals at 0.7x the measured capacity ({rate:.1f} jobs/s)",
        "jobs": s["jobs"], "direct_jobs": n_direct, "barnes_hut_jobs": count - n_direct,
        "steps_per_s": s["steps_per_s"], "unit": "job-steps/s",
        "jobs_per_s": s["jobs_per_s"],
        "body_steps_per_s": s["body_steps_per_s"],
        "direct_interactions_per_s": s["interactions_per_s"],
        "latency_p50_s": s["latency_p50_s"], "latency_p99_s": s["latency_p99_s"],
        "capacity_jobs_per_s": capacity, "arrival_rate_jobs_per_s": rate,
        "saturated_steps_per_s": sat.job_steps / sat.wall_s,
        "max_concurrent_jobs": s["max_concurrent_jobs"],
        "peak_device_bytes": s["peak_device_bytes"],
        "bh_leaf_by_size_class": leaves,
        "admission": "per-class queues (direct sum, Barnes-Hut), budget in arrival order",
        "fcfs_same_arrivals": {"steps_per_s": fcfs["steps_per_s"], "latency_p50_s": fcfs["latency_p50_s"],
                               "latency_p99_s": fcfs["latency_p99_s"]},
        "bh_contexts_same_arrivals": {"steps_per_s": ctx_rep["steps_per_s"],
                                      "latency_p50_s": ctx_rep["latency_p50_s"],
                                      "latency_p99_s": ctx_rep["latency_p99_s"],
                                      "saturated_steps_per_s": ctx_sat.job_steps / ctx_sat.wall_s,
                                      "what": "Barnes-Hut jobs each on a context of a pool of 8 "
                                              "(no Barnes-Hut batch), same jobs and arrivals"},
        "sequential_steps_per_s": seq.job_steps / seq.wall_s,
        "sequential_wall_s": seq.wall_s,
        "speedup_vs_sequential": (sat.job_steps / sat.wall_s) / (seq.job_steps / seq.wall_s),
        "sequential": "the same 160 jobs one at a time, each alone on one context via nbhip_step "
                      "(one launch per direct-sum step with the single-system launch shape; the "
                      "Barnes-Hut build and walk per step), all queued at t=0; compared with the "
                      "server's saturated run of the same jobs",
    }


def run_bh_batch(args) -> dict:
    """The served workload's Barnes-Hut jobs (seed 2026: Plummer or disk+bulge, N
    log-uniform in [16384, 262144], theta 0.5, leaf 16, eps 0.01, roots found from the
    bodies), drawn on the device into one nbhip_bh_batch: 10 steps replayed from a captured
    graph (wall time to idle), the build/walk split of 10 eager steps (HIP events), and the
    interactions of one counted pass; against the same systems one context at a time, each
    replaying 10 steps from its own graph (HIP events per system, summed)."""
    from nbody_amd import _lib, serve
    jobs = [j for j in serve.make_jobs(160, seed=SEED, device_ics=True) if j.mode == _lib.MODE_NLOGN]
    bodies = sum(j.n for j in jobs)
    steps = 10
    ctx = _lib.Context(0)
    with _lib.BhBatch(ctx) as b:
        ids = [b.add_generated(j.ics, j.n, j.params) for j in jobs]
        b.step(1, _lib.BATCH_ACCELERATE | _lib.BATCH_COUNT)
        inter = float(sum(int(b.visits(i).sum()) for i in ids))
        b.step(steps, _lib.BATCH_GRAPH)   # capture, and a warm run
        b.download(ids[0])
        t0 = time.perf_counter()
        b.step(steps, _lib.BATCH_GRAPH)
        while not b.idle():
            time.sleep(20e-6)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        build_ms, walk_ms = b.profile(steps)
        info = b.info()
    seq_ms = 0.0
    for j in jobs:
        ctx.generate(j.ics, j.n)
        ctx.step_graph(j.params, steps)
        ctx.synchronize()
        ctx.timer_start()
        ctx.step_graph(j.params, steps)
        t, _ = ctx.timer_stop()
        seq_ms += t / steps
    ctx.close()
    tflops = inter * FLOPS_PER_INTERACTION / (walk_ms * 1e-3) / 1e12
    return {
        "workload": f"{len(jobs)} Barnes-Hut systems of the served workload, {bodies} bodies in all, "
                    "theta 0.5, leaf 16, roots found from the bodies; one Barnes-Hut batch, steps "
                    "replayed from a captured graph",
        "systems": len(jobs), "bodies": bodies,
        "ms_per_step": ms, "ns_per_body_step": ms * 1e6 / bodies,
        "build_ms": build_ms, "walk_ms": walk_ms,
        "interactions_per_step": inter,
        "walk_roofline": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": tflops / FP32_PEAK_TFLOPS,
                          "flops": f"{FLOPS_PER_INTERACTION} per body-node interaction"},
        "device_bytes": int(info.device_bytes),
        "one_context_each_ms_per_step": seq_ms,
        "speedup_vs_one_context_each": seq_ms / ms,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--bh-bodies", type=int, default=1 << 20,
                    help="bodies of the Barnes-Hut leg (the C3 workload is 2^20; smaller sizes are "
                         "the served jobs' range)")
    ap.add_argument("--only", choices=["direct", "c1", "batch", "integrate", "bh", "bh_batch", "served"],
                    default=None,
                    help="run one workload (the headline fields then describe it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the multi-GPU "
              "sharding of this solver is designed but not built",
              file=sys.stderr)
        return 2
    if args.only in (None, "direct"):
        out = run_direct(args)
    else:
        out = {"metric": METRIC, "n_gpus": 1}
    if args.only in (None, "c1"):
        out["c1"] = run_c1(args)
    if args.only in (None, "batch"):
        out["batched"] = run_batch(args)
    if args.only in (None, "integrate"):
        out["integrate"] = run_integrate(args)
    if args.only in (None, "bh"):
        out["barnes_hut"] = run_bh(args)
    if args.only in (None, "bh_batch"):
        out["bh_batched"] = run_bh_batch(args)
    if args.only in (None, "served"):
        out["served"] = run_served(args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
