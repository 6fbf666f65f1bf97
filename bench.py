#!/usr/bin/env python3
"""Benchmark of the ridge-detection hot path on one MI355X.

Headline metric: evolve pair-evaluations per second at 1M 3-D points (the helix workload:
t ~ U[0, 10 pi], (cos t, sin t, 0.15 t) + N(0, 0.05^2), r = 0.2, FP32).  A step is one
complete evolve iteration over the cloud: nearest-chosen scan, ordering of the members,
ordered centroid sums and the shift, i.e. N x K pair evaluations with K chosen points.
The chosen set is the greedy choose at r on the same cloud.  Inputs are resident in HBM
when the timed region starts.

Also reported on the same JSON line:
  roofline     -- the scan kernel (dominant) against the FP32 VALU peak, with its average
                  duration measured by HIP events on the library's stream;
  cpu_baseline -- the reference CPU evolve pass (OpenMP build of the reference headers, or
                  the oracle's OpenMP restatement when that build is absent) on the host;
  secondary    -- per-iteration device time; the 512-cloud batched run against a per-cloud
                  loop; forward + backward time of the differentiated evolve.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--dtype f32|f64] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 vector peak (dense)
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    """Host cores this process may use and the CPU model (the CPU baseline's machine)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": cores, "cpu_model": model}


def cpu_baseline(P, S, r, budget_s=24.0):
    """Times the reference's OpenMP evolve pass (calculateClosestSphereCenter_omp +
    shiftTowardMassCenter_omp) over the same cloud and chosen set."""
    from oracle.oracle import Oracle, available
    kind = "reference" if available("ref_omp") else "port"
    o = Oracle("ref_omp" if kind == "reference" else "oracle_omp")
    threads = int(o.lib.ref_openmp_threads() if kind == "reference" else o.lib.oracle_num_threads())
    n, k = len(P), len(S)
    # bounded sample: as many points as fit the time budget, estimated from a small pass
    m = min(n, 20000)
    t0 = time.perf_counter()
    o.evolve_omp_iter(P[:m], S, r)
    rate = m * k / max(time.perf_counter() - t0, 1e-6)
    m = int(min(n, max(20000, rate * budget_s / 3)))
    reps = 0
    t0 = time.perf_counter()
    while True:
        o.evolve_omp_iter(P[:m], S, r)
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s / 2 or reps >= 25:
            break
    value = reps * m * k / el
    return {"value": value, "unit": "pair-evals/s", "cores": threads, "kind": kind,
            "sample": f"{reps} evolve pass(es) of {m} of the {n} points x all {k} chosen points, "
                      f"{el:.1f} s on {threads} host threads", "host": host_info()}


def cpu_per_config():
    """Whole ridge detection on the host (the reference's OpenMP build, all host threads) next
    to the GPU on the same inputs: the 20k-point spiral (r 0.45), a 50k-point sample of the 1M
    helix (its first 50k points, r 0.2) with the GPU also on the whole cloud, and the helix
    tilings (8 x 8 x 8 tiles split at 10 000 points) with the reference's own (whole, build,
    rdTiles, refinement) times."""
    from oracle.oracle import Oracle, available
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context
    if not available("ref_omp"):
        return None
    o = Oracle("ref_omp")
    threads = int(o.lib.ref_openmp_threads())
    out = {"kind": "reference", "threads": threads, "host": host_info()}

    def gpu_rd(P, r1, r2, cap=0):
        with Context(P.shape[1], P.dtype) as c:
            c.set_points(P)
            c.ridge_detection(r1, r2, max_evolve_iters=cap)      # warm (kernels, allocations)
            c.set_points(P)
            t0 = time.perf_counter()
            S = c.ridge_detection(r1, r2, max_evolve_iters=cap)
            el = time.perf_counter() - t0
            gpu_rd.last = S
            return el, len(S)

    P = synth.spiral2d(20000, seed=1)
    t0 = time.perf_counter()
    S, _ = o.ridge_detection(P, 0.45, 0.9, omp=True)
    cpu = time.perf_counter() - t0
    gpu, k = gpu_rd(P, 0.45, 0.9, cap=100)
    out["spiral20k_brute"] = {"cpu_s": cpu, "gpu_s": gpu, "cpu_ridge_points": len(S), "gpu_ridge_points": k,
                              "gpu_over_cpu_speedup": cpu / gpu}
    log(f"bench: CPU spiral20k whole RD {cpu:.2f} s, GPU {gpu * 1e3:.1f} ms")
    H = synth.helix3d(1_000_000, seed=2024, sigma=0.05, dtype=np.float32)
    sub = np.ascontiguousarray(H[:50000])
    t0 = time.perf_counter()
    S, _ = o.ridge_detection(sub, 0.2, 0.4, omp=True)
    cpu = time.perf_counter() - t0
    gpu, k = gpu_rd(sub, 0.2, 0.4)
    gfull, kfull = gpu_rd(H, 0.2, 0.4)
    ideal = synth.helix_curve(np.linspace(0.0, 2.0 * math.pi * 5.0, 100_000))
    R = gpu_rd.last.astype(np.float64)
    with Context(3, np.float64) as q:
        quality = {"hausdorff_to_pattern": q.hausdorff(R, ideal),
                   "median_distance_to_pattern": float(q.distance_stats(R, ideal)[0]),
                   "pattern_to_ridge_max": float(q.distance_stats(ideal, R)[3]), "r1": 0.2}
    out["helix_brute"] = {"sample": "first 50 000 of the 1M helix points", "cpu_s": cpu, "gpu_s": gpu,
                          "cpu_ridge_points": len(S), "gpu_ridge_points": k, "gpu_over_cpu_speedup": cpu / gpu,
                          "gpu_whole_1M_s": gfull, "gpu_whole_1M_ridge_points": kfull,
                          "gpu_whole_1M_quality": quality}
    log(f"bench: CPU helix 50k whole RD {cpu:.2f} s, GPU {gpu * 1e3:.1f} ms (1M: {gfull:.2f} s)")
    for alg in (0, 5):
        t0 = time.perf_counter()
        S = o.tiled(alg, H, 0.2, 0.4, 10000, [8, 8, 8], threads=threads)
        cpu = time.perf_counter() - t0
        whole, build, tiles, refine = (float(x) for x in o.last_times_ms)
        with Context(3, np.float32) as c:
            c.set_points(H)
            c.tiled(alg, 0.2, 0.4, 10000, [8, 8, 8])                # warm
            t0 = time.perf_counter()
            G = c.tiled(alg, 0.2, 0.4, 10000, [8, 8, 8])
            gpu = time.perf_counter() - t0
        out[f"helix_tiled/{TILED_NAMES[alg]}"] = {
            "cpu_s": cpu, "cpu_whole_ms": whole, "cpu_build_ms": build, "cpu_rd_tiles_ms": tiles,
            "cpu_refinement_ms": refine, "gpu_s": gpu, "cpu_ridge_points": len(S), "gpu_ridge_points": len(G),
            "gpu_over_cpu_speedup": cpu / gpu}
        log(f"bench: CPU helix tiled {alg}: {cpu:.2f} s, GPU {gpu * 1e3:.1f} ms")
    return out


def batched_vs_looped(count=512, loop_every=1, seed=7, graph=True, batch_only=False):
    """512 clouds (256 with d=2, 256 with d=3), sizes log-uniform in [2000, 500000], each
    a noisy spiral / helix / segment chain with r1 = 4 sigma, r2 = 2 r1: one batch of all
    clouds (one launch per phase; the 2-D clouds run zero-padded to 3 coordinates, which
    leaves their results bit-identical) against a per-cloud loop over every cloud (every
    `loop_every`-th with loop_every > 1), each in a context of its own dimension.  Both forms start from
    host arrays (uploads included).  The batch split by dimension (two batches, one per
    context dimension) is timed too."""
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context
    log(f"bench: generating {count} clouds")
    clouds = synth.batched_clouds(count=count, seed=seed)
    total_pts = sum(len(p) for p, _, _ in clouds)
    by_dim = {}
    for i, (p, r1, r2) in enumerate(clouds):
        by_dim.setdefault(p.shape[1], []).append(i)
    dmax = max(by_dim)
    ctxs = {d: Context(d, np.float32) for d in by_dim}
    if not graph:
        from ridgedetection_amd._lib import OPT_GRAPH
        for c in ctxs.values():
            c.set_option(OPT_GRAPH, 0)

    def run(ctx, idx):
        res = ctx.batch([clouds[i][0] for i in idx], [clouds[i][1] for i in idx], [clouds[i][2] for i in idx])
        return sum(len(r) for r in res)

    # warm-up (kernel loading, allocations) on a few small clouds of each dimension
    for d, idx in by_dim.items():
        run(ctxs[d], sorted(idx, key=lambda i: len(clouds[i][0]))[:4])
    run(ctxs[dmax], [sorted(idx, key=lambda i: len(clouds[i][0]))[0] for idx in by_dim.values()])
    log("bench: batched run (one mixed-dimension batch)")
    t0 = time.perf_counter()
    out_counts = run(ctxs[dmax], list(range(count)))
    t_batch = time.perf_counter() - t0
    tb = ctxs[dmax].timers()
    shape = {"samples_per_lane": tb["scan_ppt"], "threads_per_block": tb["scan_threads"],
             "persistent": bool(tb["scan_persistent"]), "tune_ms": tb["tune_ms"],
             "screen": SCREEN_FORMS[tb["screen_form"]],
             "member_ordering": "per-segment scatter" if tb["chain_path"] else "sort"}
    log(f"bench: batched {count} clouds in {t_batch:.2f} s")
    if batch_only:
        for c in ctxs.values():
            c.close()
        return {"clouds": count, "points": total_pts, "ridge_points": out_counts, "batched_s": t_batch,
                "scan_shape": shape}
    t0 = time.perf_counter()
    split_counts = sum(run(ctxs[d], idx) for d, idx in by_dim.items())
    t_split = time.perf_counter() - t0
    log(f"bench: batched by dimension in {t_split:.2f} s")
    if split_counts != out_counts:
        raise RuntimeError("mixed and per-dimension batches disagree")
    sample = list(range(0, count, loop_every))
    for d in ctxs:       # warm single-cloud paths
        i = by_dim[d][0]
        ctxs[d].set_points(clouds[i][0])
        ctxs[d].ridge_detection(clouds[i][1], clouds[i][2])
    t0 = time.perf_counter()
    for i in sample:
        p, r1, r2 = clouds[i]
        c = ctxs[p.shape[1]]
        c.set_points(p)
        c.ridge_detection(r1, r2)
    t_loop = time.perf_counter() - t0
    log(f"bench: looped {len(sample)} clouds in {t_loop:.2f} s")
    for c in ctxs.values():
        c.close()
    batched_rate = count / t_batch
    loop_rate = len(sample) / t_loop
    return {"clouds": count, "points": total_pts, "ridge_points": out_counts,
            "batched_clouds_per_s": batched_rate, "looped_clouds_per_s": loop_rate,
            "batched_over_looped": batched_rate / loop_rate,
            "looped_sample": (f"all {len(sample)} clouds" if loop_every == 1 else
                              f"every {loop_every}th cloud ({len(sample)} clouds, same size distribution)"),
            "batched_s": t_batch, "batched_by_dimension_s": t_split, "looped_s": t_loop, "scan_shape": shape}


def batched_device_generated(count=512, seed=7):
    """The batched workload's clouds (same sizes, kinds, noise fractions and radii rule as
    synth.batched_clouds, drawn from the same seeded sequence) generated on the device as
    the resident batch, then run: no host generation and no upload.  Extents (for sigma) come
    from the noise-free curves' closed forms / polyline vertices."""
    import math
    from ridgedetection_amd import _lib as L
    from ridgedetection_amd.synth import _rng
    g = _rng(seed)
    kinds, dims, sizes, params, sigmas, seeds, r1 = [], [], [], [], [], [], []
    for i in range(count):
        dim = 2 if i < count // 2 else 3
        n = int(round(math.exp(g.uniform(math.log(2000), math.log(500000)))))
        kind = int(g.integers(0, 3))
        frac = float(g.uniform(0.01, 0.1))
        s = int(g.integers(0, 2 ** 31))
        if kind == 0 and dim == 2:      # spiral t in [0, 6 pi]: (0.5 t cos t, 0.5 t sin t)
            kinds.append(L.GEN_SPIRAL2D)
            params.append([6 * math.pi, 0.5])
            t = np.linspace(0, 6 * math.pi, 20001)
            ext = float(np.max(np.ptp(np.stack([0.5 * t * np.cos(t), 0.5 * t * np.sin(t)]), axis=1)))
        elif kind in (0, 1):            # helix, 5 turns, pitch 0.15 (first dim coordinates)
            kinds.append(L.GEN_HELIX)
            params.append([5.0, 0.15])
            ext = 2.0 if dim == 2 else max(2.0, 0.15 * 10 * math.pi)
        else:                           # polyline through 5 vertices in [-1, 1]^dim
            kinds.append(L.GEN_POLYLINE)
            v = _rng(s).uniform(-1.0, 1.0, (5, dim))
            params.append([4] + v.ravel().tolist())
            ext = float(np.max(np.ptp(v, axis=0)))
        sigma = frac * ext
        dims.append(dim)
        sizes.append(n)
        sigmas.append(sigma)
        seeds.append(s)
        r1.append(4.0 * sigma)
    r2 = [2.0 * r for r in r1]
    with L.Context(3, np.float32) as c:
        c.batch_generate(kinds[:8], dims[:8], [min(x, 4000) for x in sizes[:8]], params[:8], sigmas[:8], seeds[:8])
        c.batch_run(r1[:8], r2[:8])     # warm-up
        log("bench: device-generated batch")
        t0 = time.perf_counter()
        c.batch_generate(kinds, dims, sizes, params, sigmas, seeds)
        t1 = time.perf_counter()
        res = c.batch_run(r1, r2)
        t2 = time.perf_counter()
    log(f"bench: device-generated {count} clouds: generate {t1 - t0:.3f} s, run {t2 - t1:.2f} s")
    return {"clouds": count, "points": int(sum(sizes)), "ridge_points": int(sum(len(x) for x in res)),
            "generate_s": t1 - t0, "run_s": t2 - t1, "clouds_per_s": count / (t2 - t0)}


def vjp_helix(n=200_000, iters=20, every=4, r=0.2, reps=3):
    """Differentiated evolve on the 200k-point helix (sigma 0.05, r 0.2): forward pass of
    `iters` iterations keeping checkpoints every `every` iterations, then the backward pass
    (recomputation of the earlier runs + the VJP kernels) for the helix loss gradient.
    Times are the library's own wall clock (device buffers; the host staging of the gradient
    arrays is excluded), against the forward and against plain evolve iterations (no
    checkpoints, evolve_fixed, timed by the same library clock)."""
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context
    P = synth.helix3d(n, seed=2024, dtype=np.float32)
    ctx = Context(3, np.float32)
    ctx.set_points(P)
    ctx.choose(r)
    ctx.evolve_fwd(r, iters, every)           # warm-up (allocations, kernel loading)
    K = ctx.count()
    _, G = synth.helix_loss(ctx.points())
    G = G.astype(np.float32)
    ctx.evolve_bwd(G)
    ctx.choose(r)
    ctx.evolve_fixed(r, iters)                 # warm-up of the plain form
    fwd, bwd, plain = [], [], []
    for _ in range(reps):
        ctx.choose(r)
        t = ctx.timers()
        ctx.evolve_fwd(r, iters, every)
        ctx.evolve_bwd(G)
        t2 = ctx.timers()
        fwd.append(t2["forward_ms"] - t["forward_ms"])
        bwd.append(t2["backward_ms"] - t["backward_ms"])
        ctx.choose(r)
        t = ctx.timers()
        ctx.evolve_fixed(r, iters)
        plain.append(ctx.timers()["evolve_ms"] - t["evolve_ms"])   # the library's clock, as forward_ms
    ctx.close()
    f, b, p = (float(np.median(x)) for x in (fwd, bwd, plain))
    return {"points": n, "chosen": K, "iterations": iters, "checkpoint_every": every,
            "forward_ms": f, "backward_ms": b, "plain_evolve_ms": p,
            "fwd_plus_bwd_over_fwd": (f + b) / f, "fwd_plus_bwd_over_plain_evolve": (f + b) / p}


def general_path(P, r, steps):
    """The same iterations through the general ordering path (stable radix sort of the labels +
    chain ranges, RDHIP_OPT_CHAIN = 0) that batches and tiles use: ms per iteration, graph
    replay, autotuned scan shape."""
    from ridgedetection_amd._lib import Context, OPT_CHAIN
    with Context(P.shape[1], P.dtype) as c:
        c.set_option(OPT_CHAIN, 0)
        c.set_points(P)
        c.choose(r)
        c.evolve_fixed(r, 8)
        t0 = time.perf_counter()
        c.evolve_fixed(r, steps)
        el = time.perf_counter() - t0
    return {"ms_per_step": el * 1e3 / steps, "pair_evals_per_s": None}


SCREEN_FORMS = {0: "direct", 1: "expanded", 2: "expanded on the matrix cores"}


def screen_forms(P, r, steps):
    """The same iterations with each evolve-scan screen forced (RDHIP_OPT_SCREEN 0 direct, 1
    expanded, 3 expanded on the matrix cores; the scan shape autotuned within the form): ms per
    iteration, graph replay."""
    from ridgedetection_amd._lib import Context, OPT_SCREEN
    out = {}
    for opt, form in ((0, 0), (1, 1), (3, 2)):
        with Context(P.shape[1], P.dtype) as c:
            c.set_option(OPT_SCREEN, opt)
            c.set_points(P)
            c.choose(r)
            c.evolve_fixed(r, 8)
            t0 = time.perf_counter()
            c.evolve_fixed(r, steps)
            el = time.perf_counter() - t0
            t = c.timers()
        out[SCREEN_FORMS[form]] = {"ms_per_step": el * 1e3 / steps,
                                   "scan_shape": {"samples_per_lane": t["scan_ppt"], "threads_per_block": t["scan_threads"]}}
    return out


def fp64_helix(n=1_000_000, steps=40, warmup=8, r=0.2):
    """The headline workload in FP64 (same seeded helix, cast): scan kernel and whole
    iteration times, against the FP64 vector peak."""
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context, OPT_PROFILE
    P = synth.helix3d(n, seed=2024, sigma=0.05, turns=5.0, pitch=0.15, dtype=np.float64)
    with Context(3, np.float64) as c:
        c.set_points(P)
        S0, _ = c.choose(r)
        K = len(S0)
        c.evolve_fixed(r, warmup)
        t0 = time.perf_counter()
        c.evolve_fixed(r, steps)                 # graph replay, as the headline
        el = time.perf_counter() - t0
        c.set_option(OPT_PROFILE, 1)
        c.evolve_fixed(r, 2)
        tb = c.timers()
        c.evolve_fixed(r, steps)                 # one by one, HIP events around the kernels
        ta = c.timers()
    launches = max(ta["scan_launches"] - tb["scan_launches"], 1)
    scan_ms = (ta["scan_kernel_ms"] - tb["scan_kernel_ms"]) / launches
    iter_ms = (ta["iteration_ms"] - tb["iteration_ms"]) / launches
    flops = n * K * 9 + n + 3 * n
    return {"points": n, "chosen": K, "pair_evals_per_s": n * K * steps / el, "ms_per_step": el * 1e3 / steps,
            "scan_kernel_ms": scan_ms, "evolve_iteration_device_ms": iter_ms,
            "scan_TFLOPs": flops / (scan_ms * 1e-3) / 1e12,
            "scan_frac_of_fp64_peak": flops / (scan_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "scan_shape": {"samples_per_lane": ta["scan_ppt"], "threads_per_block": ta["scan_threads"],
                           "screen": SCREEN_FORMS[ta["screen_form"]]}}


def fp32_vs_fp64():
    """The error of FP32 results against FP64 on the same coordinates.  End to end (the ridge
    is a discontinuous function of the input, so the two can differ by a flipped near tie):
    ridge point counts and the Hausdorff distance between the FP32 and FP64 ridges of the 20k
    spiral.  At frozen memberships (rounding only): 20 FP32 evolve iterations of the 200k
    helix on the GPU against the same iterations replayed in FP64 (numpy) with the GPU's labels
    and moved coordinates, max |S32 - S64| / extent."""
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context
    out = {}
    P = synth.spiral2d(20000, seed=1)
    res = {}
    for dt in (np.float32, np.float64):
        with Context(2, dt) as c:
            c.set_points(P.astype(dt))
            res[dt] = c.ridge_detection(0.45, 0.9, max_evolve_iters=100)
    with Context(2, np.float64) as c:
        h = c.hausdorff(res[np.float32].astype(np.float64), res[np.float64])
    ext = float(np.max(P.max(axis=0) - P.min(axis=0)))
    out["spiral20k"] = {"fp32_points": len(res[np.float32]), "fp64_points": len(res[np.float64]),
                        "hausdorff": h, "hausdorff_over_extent": h / ext}
    P = synth.helix3d(200000, seed=11)
    r, T = 0.2, 20
    with Context(3, np.float32) as c:
        c.set_points(P)
        S0, _ = c.choose(r)
        S = S0
        labels, moved = [], []
        for _ in range(T):
            S, _, _, _, lab, mv = c.evolve_iteration(r)
            labels.append(lab)
            moved.append(mv)
    P64 = P.astype(np.float64)
    S64 = S0.astype(np.float64)
    K, d = S64.shape
    for lab, mv in zip(labels, moved):
        inside = lab >= 0
        cnt = np.bincount(lab[inside], minlength=K).astype(np.float64)
        for j in range(d):
            sums = np.bincount(lab[inside], weights=P64[inside, j], minlength=K)
            upd = ((mv >> j) & 1).astype(bool) & (cnt > 0)
            S64[upd, j] = sums[upd] / cnt[upd]
    ext = float(np.max(P.max(axis=0) - P.min(axis=0)))
    out["helix200k_frozen_memberships"] = {"iterations": T, "chosen": K,
                                           "max_abs_err_over_extent": float(np.max(np.abs(S - S64))) / ext,
                                           "bound": 1e-5}
    return out


def tolerance_mode():
    """Tolerance mode (parallel sums in a wider type, RDHIP_OPT_SUM_MODE = 1) against the exact
    mode on the workloads whose wall time sits on the serial-sum floor: the 512-cloud batch and
    the 6-D Bezier grouped tiling.  Times, ridge point counts and the Hausdorff distance of each
    tolerance-mode ridge to the exact one (relative to the evolve radius)."""
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context, OPT_SUM_MODE
    out = {}
    clouds = synth.batched_clouds(count=512, seed=7)
    dmax = max(p.shape[1] for p, _, _ in clouds)
    res, secs = {}, {}
    for mode in (0, 1):
        with Context(dmax, np.float32) as c:
            c.set_option(OPT_SUM_MODE, mode)
            small = sorted(range(len(clouds)), key=lambda i: len(clouds[i][0]))[:4]
            c.batch([clouds[i][0] for i in small], [clouds[i][1] for i in small], [clouds[i][2] for i in small])
            t0 = time.perf_counter()
            res[mode] = c.batch([p for p, _, _ in clouds], [a for _, a, _ in clouds], [b for _, _, b in clouds])
            secs[mode] = time.perf_counter() - t0
        log(f"bench: batch sum mode {mode}: {secs[mode]:.2f} s")
    rel = []
    with Context(3, np.float64) as c:
        for (p, r1, _), ex, to in zip(clouds, res[0], res[1]):
            if len(ex) and len(to):
                a = np.zeros((len(to), 3)); a[:, :p.shape[1]] = to
                b = np.zeros((len(ex), 3)); b[:, :p.shape[1]] = ex
                rel.append(c.hausdorff(a, b) / r1)
    out["batched_512_clouds"] = {
        "exact_s": secs[0], "tolerance_s": secs[1], "speedup": secs[0] / secs[1],
        "exact_clouds_per_s": 512 / secs[0], "tolerance_clouds_per_s": 512 / secs[1],
        "ridge_points_exact": int(sum(len(x) for x in res[0])), "ridge_points_tolerance": int(sum(len(x) for x in res[1])),
        "clouds_with_equal_counts": int(sum(len(a) == len(b) for a, b in zip(res[0], res[1]))),
        "hausdorff_over_r1_median": float(np.median(rel)), "hausdorff_over_r1_max": float(np.max(rel))}
    P = synth.bezier_union(2_000_000, dim=6, seed=6, dtype=np.float32)
    tres, tt = {}, {}
    for mode in (0, 1):
        with Context(6, np.float32) as c:
            c.set_option(OPT_SUM_MODE, mode)
            c.set_points(P)
            t0 = time.perf_counter()
            tres[mode] = c.tiled(1, 0.08, 0.16, 10000, [4] * 6)
            tt[mode] = (time.perf_counter() - t0, c.timers()["evolve_ms"])
        log(f"bench: bezier6 grouped sum mode {mode}: {tt[mode][0]:.2f} s")
    with Context(6, np.float64) as c:
        h = c.hausdorff(tres[1].astype(np.float64), tres[0].astype(np.float64))
    out["bezier6/grouped tree, mixed RD"] = {
        "exact_wall_s": tt[0][0], "tolerance_wall_s": tt[1][0], "exact_evolve_ms": tt[0][1],
        "tolerance_evolve_ms": tt[1][1], "ridge_points_exact": len(tres[0]), "ridge_points_tolerance": len(tres[1]),
        "hausdorff_over_r1": h / 0.08}
    return out


TILED_NAMES = {0: "local tree, mixed RD", 1: "grouped tree, mixed RD", 2: "local tree, local RD",
               4: "grouped tree, local RD", 5: "extended tiles, mixed RD", 6: "extended tiles, local RD"}


def tiled_runs(workloads=("helix", "bezier6"), algorithms=(6, 1, 0), graph=True):
    """Tiled ridge detection on the helix (1M 3-D points, sigma 0.05, r1 0.2, r2 0.4, 8 tiles
    per dimension, tiles split at 10 000 samples) and on the 6-D Bezier union (2M points, 4
    curves, noise 0.02, 5% outliers, r1 0.08, r2 0.16, 4 tiles per dimension).  For each
    algorithm: wall time, the phases, and the evolve iterations' algorithmic bytes (every
    tile's own + neighbour samples read once per iteration, chosen points read, summed and
    updated) over their time, against the 8 TB/s HBM peak.  Quality against the noise-free
    pattern (the reference's assessment_quality measures, on the device): the Hausdorff
    distance, the median and largest distance from a ridge point to the pattern, the fraction
    of ridge points within r1 of it (the rest mark outliers), and the median and largest
    distance from a pattern point to the ridge (how well the ridge covers the pattern)."""
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context, OPT_GRAPH
    out = {}
    for w in workloads:
        if w == "helix":
            P = synth.helix3d(1_000_000, seed=2024, sigma=0.05, dtype=np.float32)
            r1, r2, tpd, cap = 0.2, 0.4, [8, 8, 8], 10000
            ideal = synth.helix_curve(np.linspace(0.0, 2.0 * math.pi * 5.0, 100_000))
        else:
            P = synth.bezier_union(2_000_000, dim=6, seed=6, dtype=np.float32)
            r1, r2, tpd, cap = 0.08, 0.16, [4] * 6, 10000
            ideal = synth.bezier_union_curves(dim=6, curves=4, seed=6)
        for alg in algorithms:
            ctx = Context(P.shape[1], np.float32)
            if not graph:
                ctx.set_option(OPT_GRAPH, 0)
            ctx.set_points(P)
            log(f"bench: tiled {w} algorithm {alg}")
            t0 = time.perf_counter()
            res = ctx.tiled(alg, r1, r2, cap, tpd)
            wall = time.perf_counter() - t0
            t = ctx.timers()
            with Context(P.shape[1], np.float64) as q:
                h = q.hausdorff(res.astype(np.float64), ideal)
                st = q.distance_stats(res.astype(np.float64), ideal)
                near = q.set_distances(res.astype(np.float64), ideal) <= r1
                cov = q.distance_stats(ideal, res.astype(np.float64))
            ctx.close()
            sec = t["evolve_ms"] * 1e-3
            gbs = t["evolve_bytes"] / sec / 1e9 if sec > 0 else 0.0
            pps = t["evolve_pairs"] / sec if sec > 0 else 0.0
            tflops = pps * 3 * P.shape[1] / 1e12        # the reference's 3d flops per pair
            out[f"{w}/{TILED_NAMES[alg]}"] = {
                "points": len(P), "dim": P.shape[1], "ridge_points": len(res), "wall_s": wall,
                "quality": {"hausdorff_to_pattern": h, "median_distance_to_pattern": float(st[0]),
                            "max_distance_to_pattern": float(st[3]), "r1": r1,
                            "fraction_within_r1_of_pattern": float(near.mean()) if len(near) else 0.0,
                            "pattern_to_ridge_median": float(cov[0]), "pattern_to_ridge_max": float(cov[3])},
                "build_ms": t["build_ms"], "choose_ms": t["choose_ms"], "evolve_ms": t["evolve_ms"],
                "decimate_ms": t["decimate_ms"], "evolve_iterations": t["evolve_iterations"],
                "rounds": t["rounds"], "evolve_GBps": gbs, "evolve_frac_of_hbm_peak": gbs / HBM_PEAK_GBS,
                "evolve_pair_evals_per_s": pps, "evolve_TFLOPs": tflops,
                "evolve_frac_of_fp32_valu_peak": tflops / FP32_PEAK_TFLOPS,
                "scan_shape": {"samples_per_lane": t["scan_ppt"], "threads_per_block": t["scan_threads"],
                               "persistent": bool(t["scan_persistent"]), "screen": SCREEN_FORMS[t["screen_form"]]}}
    return out


def batch_cloud(idx, count=512, seed=7):
    """Cloud `idx` of the batched workload alone (brute-force ridge detection from a host
    array, eager launches so that a kernel trace sees every iteration)."""
    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context, OPT_GRAPH
    P, r1, r2 = synth.batched_clouds(count=count, seed=seed)[idx]
    with Context(P.shape[1], np.float32) as c:
        c.set_option(OPT_GRAPH, 0)
        c.set_points(P)
        t0 = time.perf_counter()
        S = c.ridge_detection(r1, r2)
        wall = time.perf_counter() - t0
        t = c.timers()
    print(json.dumps({"cloud": idx, "points": len(P), "dim": P.shape[1], "r1": r1, "ridge_points": len(S),
                      "wall_s": wall, "evolve_iterations": t["evolve_iterations"], "evolve_ms": t["evolve_ms"]}),
          flush=True)


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (the HIP runtime's copy kernel), repeated 3
    times: the known-size reference the HBM byte counters of the profiled kernels are scaled
    by (FETCH_SIZE should read `mib` MiB per copy, WRITE_SIZE the same)."""
    import ctypes as C
    from ridgedetection_amd._lib import load
    load()                                   # the library's HIP runtime
    hip = C.CDLL("libamdhip64.so.7")
    n = int(mib) << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(n), 3) != 0:    # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


def measured_traffic(n, dtype):
    """HBM bytes per scan launch from the committed rocprofv3 counter summary of this workload
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "evolve_scan_helix1M_f32_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("points") != n or t.get("dtype") != dtype:
        return None
    return t.get("traffic_bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the library replays captured batches of 8 iterations: 8 warm-up steps capture the graph,
    # and 40 timed steps are 5 whole batches
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--ppt", type=int, default=0, help="samples per lane (default: autotuned)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-configs", action="store_true",
                    help="skip the host's whole ridge detections per workload (the evolve-pass baseline stays)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-cloud batched measurement")
    ap.add_argument("--no-vjp", action="store_true", help="skip the differentiated-evolve measurement")
    ap.add_argument("--no-tiled", action="store_true", help="skip the tiled measurements")
    ap.add_argument("--tiled-case", action="append", default=[], metavar="WORKLOAD:ALG",
                    help="only run these tiled cases (helix|bezier6 : algorithm id), print them and exit")
    ap.add_argument("--no-graph", action="store_true",
                    help="tiled cases / the batch with eager launches instead of graph replay (kernel traces)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the 512-cloud batch (one mixed-dimension batch), print it and exit")
    ap.add_argument("--batch-cloud", type=int, default=-1, metavar="IDX",
                    help="only run cloud IDX of the 512-cloud batch on its own (profiling) and exit")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline iterations (profiling): no secondary measurements, no CPU baseline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.batch_cloud >= 0:
        batch_cloud(args.batch_cloud)
        return
    if args.batch_only:
        print(json.dumps(batched_vs_looped(graph=not args.no_graph, batch_only=True)), flush=True)
        return
    if args.tiled_case:
        for case in args.tiled_case:
            w, alg = case.split(":")
            print(json.dumps(tiled_runs(workloads=(w,), algorithms=(int(alg),), graph=not args.no_graph)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from ridgedetection_amd import synth
    from ridgedetection_amd._lib import Context, OPT_PROFILE, OPT_PPT

    dt = np.float32 if args.dtype == "f32" else np.float64
    r = 0.2
    d = 3
    log(f"bench: generating helix n={args.n} {args.dtype}")
    P = synth.helix3d(args.n, seed=2024, sigma=0.05, turns=5.0, pitch=0.15, dtype=dt)
    ctx = Context(d, dt)
    ctx.set_points(P)
    S0, _ = ctx.choose(r)
    K = len(S0)
    log(f"bench: K={K} chosen points")
    if args.ppt:
        ctx.set_option(OPT_PPT, args.ppt)
    # timed region: the production form (iterations replayed from a captured graph); every
    # library call returns with its stream synchronised
    ctx.evolve_fixed(r, args.warmup)
    t0 = time.perf_counter()
    ctx.evolve_fixed(r, args.steps)
    el = time.perf_counter() - t0
    log(f"bench: {args.steps} iterations in {el * 1e3:.2f} ms")
    # then the same iterations launched one by one with HIP events around the scan kernel and
    # around each iteration on the library's stream (the roofline's kernel time)
    ctx.set_option(OPT_PROFILE, 1)
    ctx.evolve_fixed(r, 2)
    t_before = ctx.timers()
    ctx.evolve_fixed(r, args.steps)
    t_after = ctx.timers()
    ctx.set_option(OPT_PROFILE, 0)

    n = args.n
    pairs = n * K
    value = pairs * args.steps / el
    launches = t_after["scan_launches"] - t_before["scan_launches"]
    scan_ms = (t_after["scan_kernel_ms"] - t_before["scan_kernel_ms"]) / max(launches, 1)
    iter_ms = (t_after["iteration_ms"] - t_before["iteration_ms"]) / max(launches, 1)
    flops = n * K * 3 * d + n + n * d            # reference convention, per scan launch
    achieved = flops / (scan_ms * 1e-3) / 1e12
    peak = FP32_PEAK_TFLOPS if dt == np.float32 else FP64_PEAK_TFLOPS

    # the rest of the path, for context: full brute-force ridge detection on the cloud
    secondary = {
        "chosen_points": K,
        "scan_kernel_ms": round(scan_ms, 4),
        "evolve_iteration_device_ms": round(iter_ms, 4),
        "scan_pair_evals_per_s": pairs / (scan_ms * 1e-3),
        "scan_flops_per_launch": flops,
    }
    result = {
        "metric": "evolve pair-evals/s at 1M 3-D pts",
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded helix, resident in HBM)",
        "config": {"workload": "helix 1M 3-D points, sigma 0.05, r 0.2, brute-force evolve iteration",
                   "points": n, "dim": d, "r": r, "chosen": K,
                   "scan_shape": {"samples_per_lane": t_after["scan_ppt"], "threads_per_block": t_after["scan_threads"],
                                  "screen": SCREEN_FORMS[t_after["screen_form"]], "autotuned": not args.ppt}},
        "roofline": {"bound": "mfma" if t_after["screen_form"] == 2 else "valu", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(n, args.dtype),
                     "kernel": "evolve_scan_hist_kernel" if t_after["chain_path"] else "evolve_scan_kernel",
                     "avg_launch_ms": scan_ms,
                     "flops_per_launch": flops},
        "secondary": secondary,
    }
    ctx.close()
    if args.headline_only:
        print(json.dumps(result), flush=True)
        return
    log("bench: the same iterations through the general ordering path")
    gp = general_path(P, r, args.steps)
    gp["pair_evals_per_s"] = pairs * 1e3 / gp["ms_per_step"]
    result["secondary"]["general_ordering_path"] = gp
    log("bench: the same iterations with each screen form forced")
    result["secondary"]["screen_forms"] = screen_forms(P, r, args.steps)
    if dt == np.float32:
        log("bench: FP64 helix and FP32-vs-FP64 errors")
        result["secondary"]["fp64_helix_1M"] = fp64_helix(n=args.n)
        result["secondary"]["fp32_vs_fp64"] = fp32_vs_fp64()
    if not args.no_vjp and dt == np.float32:
        log("bench: differentiated evolve (200k helix, 20 iterations)")
        result["secondary"]["vjp_helix_200k"] = vjp_helix()
    if not args.no_tiled and dt == np.float32:
        result["secondary"]["tiled"] = tiled_runs()
    if not args.no_batch:
        result["secondary"]["batched_512_device_generated"] = batched_device_generated()
        result["secondary"]["batched_512_clouds"] = batched_vs_looped()
        if dt == np.float32 and not args.no_tiled:
            result["secondary"]["tolerance_mode"] = tolerance_mode()
    if not args.no_cpu:
        log("bench: CPU baseline")
        result["cpu_baseline"] = cpu_baseline(P, S0, r)
        if dt == np.float32 and not args.no_cpu_configs:
            log("bench: CPU whole ridge detection per workload")
            result["secondary"]["cpu_per_config"] = cpu_per_config()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
