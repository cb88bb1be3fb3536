"""Benchmark: W-projection gridding + degridding throughput on one MI355X.

Headline workload (the metric's, "batched vs plain"): 64 independent jobs,
visibilities log-uniform in [1e4, 2e6] per job, grids of 1024^2, 2048^2 or
4096^2, wSize 33, 65 or 129, baselines 1000-6000 m, complex64, tile shape
chosen by the batch autotuner.  One step = re-bin every job's visibilities on
the device + zero the grids + grid every job + degrid every job, all in one
captured hipGraph of persistent launches over a device work queue, on
device-resident inputs.  value = (grid points gridded + grid points degridded)
/ step time, beside the same jobs run as one plain launch per job.

Also measured (fields of the same JSON line): the roofline of the headline's
dominant kernel; config 1 (160 000 samples, 4096^2, width 129) grid and degrid
rates beside the published P100 figures, in FP32 and FP16 storage with max
and RMS errors against the device's FP64 reference kernels; the spectral
workload (1 048 576 samples x 8 channels); the serving loop (Poisson
requests, p50/p99 latency); the major cycle; and the CPU baseline on the
host's cores (the reference's own gridKernel/degridKernel compiled from its
sources when that checker binary was built, else the oracle's restatement),
on a bounded sample of the headline jobs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "askap-benchmarks_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published
ENGINE_CLOCK_GHZ = 2.4  # MI355X peak engine clock, published
# published GPU figures for this path: Tesla P100, CUDA d_gridKernel and
# d_degridKernel<64>, config 1 with u, v, w in [-1000, 1000] m
# (attic/tConvolveCuda/slurm-2759.out:16-37), M grid points/s
P100_GRID_MPTS = 14013.5
P100_DEGRID_MPTS = 38036.6
# published per-process rates of the current generation's tConvolveMPI on one
# Magnus node with one process (README.md:148-184), (gridding, degridding) Mpix/s,
# for runType 0 (continuum) and 1 (spectral)
MAGNUS_NP1_MPIX = {0: (294.649, 206.756), 1: (482.345, 390.47)}
# the GPU box grants one GPU's job this many host CPUs, whatever the machine has
BOX_CPU_SHARE = 16


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def alg_bytes(points, nVis, gridCells, op):
    """Algorithmic bytes of one launch: one complex64 C tap per grid point, the
    grid read (degrid) or read + written (grid) once, 20 B per visibility
    (iu, iv, cOffset, data)."""
    return 8 * points + (16 if op == "grid" else 8) * gridCells + 20 * nVis


TUNING_KEYS = ("tile_w", "tile_h", "chunk", "slice_keys", "phases", "xcd_split", "kernel",
               "degrid_kernel", "grid_phases", "grid_cols")


def traffic_record(name):
    """The `traffic-record:` JSON line of profiles/<name>.md (counter passes
    over this program), or {}."""
    try:
        lines = open(os.path.join(REPO, "profiles", name + ".md")).read().splitlines()
    except OSError:
        return {}
    rec = {}
    for ln in lines:
        if ln.startswith("traffic-record:"):
            rec = json.loads(ln.split(":", 1)[1])
    return rec


def read_ceilings(ctx):
    """Streaming read bandwidth (GB/s) of the device's levels, measured live:
    a 2 MiB buffer (held by the L2s), 64 MiB (Infinity Cache) and 1 GiB (HBM)."""
    out = {}
    for name, nbytes in (("l2", 2 << 20), ("infinity_cache", 64 << 20), ("hbm", 1 << 30)):
        buf = ctx.empty(nbytes // 4, np.uint32)
        buf.zero()
        out[name] = ctx.read_bandwidth(buf)
        buf.free()
    return out


def calibration_copy(ctx, nbytes=1 << 30, reps=2):
    """Launches the streaming copy kernel on a buffer of known size so that a
    counter run over this program also reads the counters' scale."""
    a = ctx.empty(nbytes // 4, np.uint32)
    b = ctx.empty(nbytes // 4, np.uint32)
    a.zero()
    for _ in range(reps):
        b.copy_from(a)
    ctx.sync()
    a.free()
    b.free()


def roofline(kernel, avg_ms, bytes_alg, ceilings, rec, cus):
    """Roofline of one kernel.  The C taps of these workloads are reused from
    the L2 and Infinity Cache (each slice serves hundreds to thousands of
    visibilities), so the bytes that must cross HBM are far fewer than the 8 B
    a grid point of algorithmic traffic: the binding path is L2 -> L1 (the
    CUs' vector-memory pipeline).  `achieved` = algorithmic bytes / launch
    time against the measured L2 read ceiling (MI355X publishes none); the
    HBM traffic the counters saw (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    and the L2 -> L1 bytes (TCP_TCC_READ_REQ x 128 B) stand beside it, and so
    does the share of the launch the CUs' texture address units were busy
    (TA_TA_BUSY summed over the CUs, at the peak engine clock): the unit that
    binds these kernels."""
    achieved = bytes_alg / (avg_ms * 1e-3) / 1e9
    traffic = (rec.get("bytes_per_launch") or {}).get(kernel)
    l2 = (rec.get("l2_read_bytes_per_launch") or {}).get(kernel)
    out = {"bound": "l2", "achieved": achieved, "peak": ceilings["l2"], "unit": "GB/s",
           "frac": achieved / ceilings["l2"], "traffic": traffic,
           "peak_source": "measured: tcv_read_bandwidth, 16-byte loads from every CU over a 2 MiB "
                          "buffer held by the L2s (no published L2 figure)",
           "kernel": kernel, "avg_launch_ms": avg_ms, "alg_bytes_per_launch": bytes_alg,
           "alg_bytes_per_point": 8,
           "hbm": {"traffic_bytes_per_launch": traffic,
                   "GBps": traffic / (avg_ms * 1e-3) / 1e9 if traffic else None,
                   "peak_GBps": HBM_PEAK_GBS,
                   "frac": traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None},
           "read_ceilings_GBps": ceilings}
    if l2:
        out["l2_to_l1"] = {"bytes_per_launch": l2, "GBps": l2 / (avg_ms * 1e-3) / 1e9,
                           "frac_of_l2_ceiling": l2 / (avg_ms * 1e-3) / 1e9 / ceilings["l2"],
                           "over_algorithmic": l2 / bytes_alg}
    ta = (rec.get("ta_busy_cycles_per_launch") or {}).get(kernel)
    if ta:
        out["address_unit"] = {"busy_cycles_sum": ta, "cus": cus, "clock_GHz": ENGINE_CLOCK_GHZ,
                               "busy_frac": ta / (cus * avg_ms * 1e-3 * ENGINE_CLOCK_GHZ * 1e9)}
    return out


# ---- the headline: the 64-job uneven batch (config 3) ------------------------------

def bench_batch(ctx, args, ceilings):
    from tcv import setup as S
    from tcv.gridder import BatchJobs
    t0 = time.time()
    specs = S.uneven_batch(nJobs=args.batch_jobs)
    bj = BatchJobs(ctx, specs)
    nVis = int(sum(s["nVis"] for s in specs))
    log(f"batch: {len(specs)} jobs, {nVis} visibilities, {bj.points / 1e9:.1f} G points per pass, "
        f"set up in {time.time() - t0:.1f} s")
    if args.batch_tuning:
        best, table = json.loads(args.batch_tuning), []
        bj.set_tuning(best)
    else:
        best, table = bj.autotune(reps=2)
    log(f"batch autotune: {best} (" + ", ".join(f"{t.get('tile_w')}x{t.get('tile_h')}/"
                                              f"{t.get('item_entries', 0)}/c{t.get('grid_cols', 0)} "
                                              f"{ms:.1f} ms" for t, ms in table) + ")")
    pts = 2 * bj.points

    def step():
        bj.zero_grids()
        bj.run_batched(graph=True)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    w0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        step()
    ms_step = ctx.timer_stop() / args.steps
    wall = time.perf_counter() - w0
    value = pts / (ms_step * 1e-3)
    log(f"batched: {ms_step:.1f} ms/step, {value / 1e9:.1f} G points/s (wall {wall:.2f} s)")

    # per-kernel launch times of the same batched kernels, hipEvents on the
    # library's stream (the captured graph holds no events)
    ctx.set_profiling(True)
    ctx.reset_stats()
    for _ in range(2):
        bj.zero_grids()
        bj.run_batched(graph=False)
    stats = {f: ctx.kernel_stats(f) for f in ("plan", "grid_batch", "degrid_batch")}
    ctx.set_profiling(False)
    kern = {k: {"launches": v[0], "avg_ms": v[1] / v[0] if v[0] else None} for k, v in stats.items()}
    dom = max(("grid_batch", "degrid_batch"), key=lambda k: kern[k]["avg_ms"] or 0)
    cells = int(sum(s["gSize"] ** 2 for s in specs))
    rec = traffic_record("traffic_batch_uneven64")
    if rec.get("tuning") != best:        # counters measured under another tuning
        rec = {}
    rf = roofline(dom, kern[dom]["avg_ms"],
                  alg_bytes(bj.points, nVis, cells, "grid" if dom == "grid_batch" else "degrid"),
                  ceilings, rec, ctx.cus)

    def timed(fn, reps):
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            bj.zero_grids()
            fn()
        return ctx.timer_stop() / reps

    if args.no_plain:                  # counter runs: the batched kernels alone
        t_tiled = t_plain = float("nan")
    else:
        timed(lambda: bj.run_tiled(), 1)
        t_tiled = timed(lambda: bj.run_tiled(), 2)
        t_plain = timed(lambda: bj.run_plain(), 1)
    widths = sorted({s["sSize"] for s in specs})
    out = {
        "ms_per_step": ms_step, "value": value, "wall_s": wall,
        "workload": (f"{len(specs)} jobs: visibilities log-uniform [1e4, 2e6], grids "
                     "{1024,2048,4096}^2, wSize {33,65,129}, baselines 1000-6000 m (seed 3), "
                     "complex64, tightest-fit cellSize; re-bin + grid + degrid in one hipGraph"),
        "jobs": len(specs), "visibilities": nVis, "grid_points_per_pass": bj.points,
        "kernel_widths": widths, "tuning": best,
        "autotune": [{"tuning": t, "ms_per_step": ms} for t, ms in table],
        "kernels": kern, "roofline": rf,
        "tiled_per_job": {"ms": t_tiled, "value": pts / (t_tiled * 1e-3)},
        "plain_per_job": {"ms": t_plain, "value": pts / (t_plain * 1e-3)},
        "speedup_batched_vs_plain": t_plain / ms_step,
        "speedup_batched_vs_tiled": t_tiled / ms_step,
    }
    log(f"batch: batched {ms_step:.1f} ms, tiled/job {t_tiled:.1f} ms, plain/job {t_plain:.1f} ms "
        f"-> batched/plain {out['speedup_batched_vs_plain']:.2f}x; dominant {dom} "
        f"{kern[dom]['avg_ms']:.1f} ms")
    return bj, out


# ---- config 1 (attic constants) and FP16 storage (config 4) ------------------------

def bench_config1(ctx, variant, reps, tuning):
    """Config 1: 160 000 samples, 1 channel, 4096^2 grid, wSize 33, width 129,
    the reference generator (variant "reference": u, v, w in [-1000, 1000] m,
    the P100 figures' inputs; "wide": [-2000, 2000] m).  Tiled grid and degrid
    timed apart, as the P100 figures are; FP32 and FP16 storage, each with max
    and RMS error against the FP64 reference kernels on the device."""
    import tcv
    from tcv import setup as S
    prob = S.attic_problem(variant)
    job = tcv.GridJob(ctx, prob, tuning=tuning)
    assert job.nBad == 0
    pts = job.points
    nv, g = prob.nVis, prob.gSize
    job.zero_grid()
    job.grid("tiled")
    job.degrid("tiled")
    ctx.sync()

    def timed(fn):
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            fn()
        return ctx.timer_stop() / reps

    t_grid = timed(lambda: job.grid("tiled"))      # accumulates: timing only
    job.zero_grid()
    job.grid("tiled")
    t_degrid = timed(lambda: job.degrid("tiled"))
    g32 = job.grid_d.download()
    o32 = job.out.download()
    # FP64 references: gridding of the same inputs, degridding of the FP32 grid
    g64d = ctx.zeros((g, g), np.complex128)
    tcv.grid_plain_f64(ctx, job.vis, job.data, job.C, g64d)
    g64 = g64d.download()
    g64d.free()
    o64d = ctx.empty(nv, np.complex128)
    tcv.degrid_plain_f64(ctx, job.vis, job.grid_d, job.C, o64d)
    o64 = o64d.download()
    o64d.free()

    def errs(a, ref):
        m = float(np.abs(ref).max())
        d = np.abs(a.astype(np.complex128) - ref)
        return {"max": float(d.max() / m), "rms": float(np.sqrt(np.mean(d * d)) / m)}

    # FP16 storage of C (per-w-plane scales) and visibilities, FP32 accumulation
    Ch = tcv.HalfTable(ctx, job.C, prob.wSize, 129 * 129 * 64)
    dh = tcv.HalfVis(ctx, job.data)
    plan16 = tcv.Plan(ctx, job.vis, tap_bytes=4)
    gh = ctx.zeros((g, g), np.complex64)
    plan16.grid_half(dh, Ch, gh)
    g16 = gh.download()
    oh = ctx.empty(nv, np.complex64)
    plan16.degrid_half(job.grid_d, Ch, oh)          # the FP32 grid: the same input as o64
    o16 = oh.download()
    t_grid16 = timed(lambda: plan16.grid_half(dh, Ch, gh))
    t_degrid16 = timed(lambda: plan16.degrid_half(job.grid_d, Ch, oh))
    # FP32 with the library's autotuner (tcv_autotune's built-in search on this job),
    # run in a context of its own so that its cached choice leaves the other
    # variant's automatic plan alone; checked against the same FP64 references
    actx = tcv.Context(ctx.device)
    try:
        sg, so = ctx.zeros((g, g), np.complex64), ctx.empty(nv, np.complex64)
        best = tcv.autotune(actx, job.vis, job.data, job.C, sg, so, op=3)
        sg.free()
        so.free()
        tuned = {k: best[k] for k in TUNING_KEYS + ("tap_bytes", "item_entries")}
    finally:
        actx.close()
    plan_a = tcv.Plan(ctx, job.vis, **tuned)
    ga, oa = ctx.zeros((g, g), np.complex64), ctx.empty(nv, np.complex64)
    plan_a.grid(job.data, job.C, ga)
    plan_a.degrid(job.grid_d, job.C, oa)
    ga_h, oa_h = ga.download(), oa.download()
    t_grid_a = timed(lambda: plan_a.grid(job.data, job.C, ga))
    t_degrid_a = timed(lambda: plan_a.degrid(job.grid_d, job.C, oa))
    ga.free()
    oa.free()
    plan_a.close()
    out = {
        "workload": (f"config 1 ({variant}): 160000 samples x 1 channel, u,v,w in "
                     f"[-{prob.meta['baseline_m'] // 2}, {prob.meta['baseline_m'] // 2}] m "
                     "(LCG next=1), data 1+0i, 4096^2 grid, wSize 33, width 129"),
        "grid_points_per_pass": pts,
        "w_planes_used": int((job.numPerPlane > 0).sum()),
        "fp32": {"grid_ms": t_grid, "degrid_ms": t_degrid,
                 "grid_Mpts": pts / (t_grid * 1e-3) / 1e6, "degrid_Mpts": pts / (t_degrid * 1e-3) / 1e6,
                 "grid_err_vs_fp64": errs(g32, g64), "degrid_err_vs_fp64": errs(o32, o64)},
        "fp16": {"grid_ms": t_grid16, "degrid_ms": t_degrid16,
                 "grid_Mpts": pts / (t_grid16 * 1e-3) / 1e6,
                 "degrid_Mpts": pts / (t_degrid16 * 1e-3) / 1e6,
                 "grid_err_vs_fp64": errs(g16, g64), "degrid_err_vs_fp64": errs(o16, o64),
                 "alg_bytes_per_point": 4},
        "fp32_autotuned": {"grid_ms": t_grid_a, "degrid_ms": t_degrid_a,
                           "grid_Mpts": pts / (t_grid_a * 1e-3) / 1e6,
                           "degrid_Mpts": pts / (t_degrid_a * 1e-3) / 1e6,
                           "vs_p100_grid": pts / (t_grid_a * 1e-3) / 1e6 / P100_GRID_MPTS,
                           "vs_p100_degrid": pts / (t_degrid_a * 1e-3) / 1e6 / P100_DEGRID_MPTS,
                           "grid_err_vs_fp64": errs(ga_h, g64), "degrid_err_vs_fp64": errs(oa_h, o64),
                           "tuning": tuned},
        "tuning": {k: job.plan.info()[k] for k in TUNING_KEYS},
    }
    f = out["fp32"]
    f["vs_p100_grid"] = f["grid_Mpts"] / P100_GRID_MPTS
    f["vs_p100_degrid"] = f["degrid_Mpts"] / P100_DEGRID_MPTS
    fa = out["fp32_autotuned"]
    log(f"config 1 {variant}: grid {f['grid_Mpts']:.0f} M pts/s ({f['vs_p100_grid']:.1f}x P100), "
        f"degrid {f['degrid_Mpts']:.0f} ({f['vs_p100_degrid']:.1f}x); fp16 grid "
        f"{out['fp16']['grid_Mpts']:.0f}, degrid {out['fp16']['degrid_Mpts']:.0f}; errors vs FP64 "
        f"fp32 {f['grid_err_vs_fp64']['max']:.1e}/{f['degrid_err_vs_fp64']['max']:.1e}, fp16 "
        f"{out['fp16']['grid_err_vs_fp64']['max']:.1e}/{out['fp16']['degrid_err_vs_fp64']['max']:.1e}; "
        f"autotuned fp32 grid {fa['grid_Mpts']:.0f} ({fa['vs_p100_grid']:.1f}x), degrid "
        f"{fa['degrid_Mpts']:.0f} ({fa['vs_p100_degrid']:.1f}x), errors "
        f"{fa['grid_err_vs_fp64']['max']:.1e}/{fa['degrid_err_vs_fp64']['max']:.1e}, phases "
        f"{tuned['grid_phases']}/{tuned['phases']}")
    plan16.close()
    job.plan.close()
    return out


# ---- config 2 (the spectral scale) ----------------------------------------------

def bench_spectral(ctx, args, tuning, steps):
    import tcv
    from tcv import setup as S
    t0 = time.time()
    prob = S.spectral_problem(nSamples=args.nsamples, nChan=args.nchan)
    job = tcv.GridJob(ctx, prob, tuning=tuning)
    assert job.nBad == 0, f"{job.nBad} visibilities out of range"
    points = job.points
    info = job.plan.info()
    log(f"spectral: {prob.nVis} visibilities, C {prob.C.nbytes / 2**30:.2f} GiB, set up in "
        f"{time.time() - t0:.1f} s; plan {info}")

    def step():
        job.plan.build()
        job.zero_grid()
        job.grid("tiled")
        job.degrid("tiled")

    step()
    ctx.sync()
    ctx.set_profiling(True)
    ctx.reset_stats()
    ctx.timer_start()
    for _ in range(steps):
        step()
    ms_step = ctx.timer_stop() / steps
    stats = {f: ctx.kernel_stats(f) for f in ("plan", "grid_tiled", "degrid_tiled")}
    ctx.set_profiling(False)
    kern = {k: {"launches": v[0], "avg_ms": v[1] / v[0] if v[0] else None} for k, v in stats.items()}
    value = 2 * points / (ms_step * 1e-3)
    out = {"workload": "spectral: 1048576 samples x 8 channels, u,v,w in [-6000,6000] m (PCG64 seed "
                       "2024), complex Gaussian data, 4096^2 grid, wSize 129, width 129",
           "ms_per_step": ms_step, "value": value, "grid_points_per_pass": points,
           "kernels": kern, "tuning": info}
    rec = traffic_record("traffic_grid_degrid_tiled_spectral")
    if rec and all(rec.get("tuning", {}).get(k) == info.get(k) for k in TUNING_KEYS):
        out["roofline"] = {k: roofline(k, kern[k]["avg_ms"],
                                       alg_bytes(points, prob.nVis, prob.gSize ** 2, k.split("_")[0]),
                                       args._ceilings, rec, ctx.cus)
                           for k in ("grid_tiled", "degrid_tiled")}
    log(f"spectral: {ms_step:.2f} ms/step, {value / 1e9:.1f} G points/s (grid "
        f"{kern['grid_tiled']['avg_ms']:.1f}, degrid {kern['degrid_tiled']['avg_ms']:.1f} ms)")
    if not args.no_plain:
        job.zero_grid()
        ctx.sync()
        ctx.timer_start()
        job.grid("plain")
        t_gp = ctx.timer_stop()
        ctx.timer_start()
        job.degrid("plain")
        t_dp = ctx.timer_stop()
        plain = 2 * points / ((t_gp + t_dp) * 1e-3)
        out["plain"] = {"value": plain, "grid_ms": t_gp, "degrid_ms": t_dp,
                        "speedup_tiled_vs_plain": value / plain}
        log(f"spectral plain: grid {t_gp:.1f} ms, degrid {t_dp:.1f} ms")
    if not args.no_half:
        Ch = tcv.HalfTable(ctx, job.C, prob.wSize, int(prob.sSize[0]) ** 2 * 64)
        dh = tcv.HalfVis(ctx, job.data)
        gh = ctx.zeros((prob.gSize, prob.gSize), np.complex64)
        oh = ctx.empty(prob.nVis, np.complex64)
        plan16 = tcv.Plan(ctx, job.vis, **dict(tuning, tap_bytes=4))
        plan16.grid_half(dh, Ch, gh)
        plan16.degrid_half(gh, Ch, oh)
        reps = 3
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            plan16.grid_half(dh, Ch, gh)
        t_gh = ctx.timer_stop() / reps
        ctx.timer_start()
        for _ in range(reps):
            plan16.degrid_half(gh, Ch, oh)
        t_dh = ctx.timer_stop() / reps
        out["half"] = {"value": 2 * points / ((t_gh + t_dh) * 1e-3), "grid_ms": t_gh, "degrid_ms": t_dh}
        log(f"spectral fp16: grid {t_gh:.1f} ms, degrid {t_dh:.1f} ms")
        plan16.close()
    job.plan.close()
    return out


# ---- the current generation's runType 0 and 1 --------------------------------------

def bench_current(ctx, reps):
    """current/tConvolveMPI's two reported runs on one GPU: runType 0 (continuum,
    6440^2 grid, 483 w-planes of 7-121 wide kernels) and 1 (spectral), the
    reference's own uvw generator, C(w) and setSort(0), each gridding and
    degridding timed apart as tConvolveMPI.cc:101-170 times them, in Mpix/s (its
    nPixelsGridded over the time) beside the published per-process rates and the
    requiredRate model (Benchmark.cc:702-740).  Kernel time with the plan held;
    the plan's re-binning of new u, v, w is timed beside it."""
    import tcv
    req = tcv.CurrentBenchmark.requiredRate()
    out = {}
    for rt in (0, 1):
        bm = tcv.CurrentBenchmark(ctx)
        bm.setRunType(rt)
        bm.init()
        npp = bm.numPerPlane.astype(np.int64)
        s = bm.sSize.astype(np.int64)
        pix = int((npp * s * s).sum())
        bm.runGrid()
        bm.runDegrid()
        ctx.sync()

        def timed(fn):
            ctx.sync()
            ctx.timer_start()
            for _ in range(reps):
                fn()
            return ctx.timer_stop() / reps

        t_plan = timed(lambda: bm._plan().build())
        t_grid = timed(bm.runGrid)
        t_degrid = timed(bm.runDegrid)
        # the library's autotuner on the same job, in a context of its own
        actx = tcv.Context(ctx.device)
        try:
            sg = ctx.zeros((bm.gSize, bm.gSize), np.complex64)
            so = ctx.empty(bm.samples.nVis, np.complex64)
            best = tcv.autotune(actx, bm.samples, bm.data, bm.C, sg, so, op=3)
            sg.free()
            so.free()
            tuned = {k: best[k] for k in TUNING_KEYS + ("tap_bytes", "item_entries")}
        finally:
            actx.close()
        plan_a = tcv.Plan(ctx, bm.samples, **tuned)
        t_grid_a = timed(lambda: plan_a.grid(bm.data, bm.C, bm.grid1))
        t_degrid_a = timed(lambda: plan_a.degrid(bm.grid1, bm.C, bm.outdata1))
        plan_a.close()
        mpix = lambda t: pix / (t * 1e-3) / 1e6
        pub = MAGNUS_NP1_MPIX[rt]
        out[f"runType{rt}"] = {
            "visibilities": int(bm.nVisibilitiesGridded()), "pixels_gridded": pix,
            "out_of_range": int(bm.nOutOfRange), "gSize": int(bm.gSize), "wSize": int(bm.wSize),
            "plan_ms": t_plan, "grid_ms": t_grid, "degrid_ms": t_degrid,
            "grid_Mpix_per_s": mpix(t_grid), "degrid_Mpix_per_s": mpix(t_degrid),
            "vs_magnus_process_grid": mpix(t_grid) / pub[0],
            "vs_magnus_process_degrid": mpix(t_degrid) / pub[1],
            "autotuned": {"grid_ms": t_grid_a, "degrid_ms": t_degrid_a,
                          "grid_Mpix_per_s": mpix(t_grid_a), "degrid_Mpix_per_s": mpix(t_degrid_a),
                          "tuning": tuned},
            "required_Mpix_per_s": req[rt] / 1e6,
            "tuning": {k: bm._plan().info()[k] for k in TUNING_KEYS}}
        log(f"current runType {rt}: grid {mpix(t_grid):.0f} Mpix/s ({mpix(t_grid) / pub[0]:.0f}x a "
            f"Magnus process), degrid {mpix(t_degrid):.0f} ({mpix(t_degrid) / pub[1]:.0f}x); autotuned "
            f"{mpix(t_grid_a):.0f} / {mpix(t_degrid_a):.0f}; plan {t_plan:.2f} ms; required "
            f"{req[rt] / 1e6:.1f} Mpix/s")
        bm.plan.close()
    return out


# ---- major cycle ----------------------------------------------------------------

def bench_major(ctx, args):
    """tMajorACC's imaging loop at its own scale (160 000 samples, 4096^2
    grid, wSize 33, 5 major cycles of 100 Hogbom iterations), device-resident:
    per-stage times summed over the cycles (each stage synchronised)."""
    from tcv.major import MajorCycle
    t0 = time.time()
    mc = MajorCycle(ctx, nSamples=args.major_samples, nMajor=args.major_cycles,
                    nMinor=args.major_minor)
    setup_s = time.time() - t0
    mc.cycle(0)                       # warm: plans, FFT plans, workspaces
    ctx.sync()
    mc = MajorCycle(ctx, nSamples=args.major_samples, nMajor=args.major_cycles,
                    nMinor=args.major_minor)
    ctx.sync()
    w0 = time.perf_counter()
    mc.run(timed=True)
    wall = time.perf_counter() - w0
    (pv, pp), (dv, dp) = mc.peaks[0]
    out = {"workload": f"tMajorACC: {args.major_samples} samples x 1 channel, 4096^2 grid, "
                       f"wSize 33, {args.major_cycles} major x {args.major_minor} minor cycles",
           "total_ms": wall * 1e3, "per_cycle_ms": wall * 1e3 / args.major_cycles,
           "stages_ms": {k: v * 1e3 for k, v in mc.times.items()},
           "setup_s": setup_s,
           "psf_peak": [pv, pp], "first_dirty_peak": [dv, dp]}
    log(f"major: {out['per_cycle_ms']:.1f} ms per cycle")
    del mc
    return out


# ---- serving (config 5) ------------------------------------------------------------

def serving_stream(srv, bj, mean_pts, capacity, rate, n, seed):
    """One Poisson stream of n requests at `rate` per second against `srv`:
    requests drawn from the resident jobs, 50% grid into a new grid, 50%
    degrid the job's resident grid.  Latency = submission to completion.  The
    queue wait of the first and last fifths of the stream tells whether it
    was steady (the server keeping up) or growing."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    ops = rng.integers(0, 2, n)
    tmpl = rng.integers(0, len(bj.jobs), n)
    t0 = srv.clock() + 0.05
    ids = []
    for i in range(n):
        target = t0 + arrivals[i]
        while True:
            d = target - srv.clock()
            if d <= 0:
                break
            time.sleep(min(d, 0.002) if d > 0.0005 else 0)
        j = bj.jobs[tmpl[i]]
        if ops[i] == 0:
            ids.append(srv.submit(0, j["vis"], j["C"], data=j["data"]))
        else:
            ids.append(srv.submit(1, j["vis"], j["C"], grid=j["grid"]))
    res = []
    for r in ids:
        res.append(srv.wait(r, timeout=900))
    bad = sum(r["status"] != 0 for r in res)
    lat = np.array([r["completed"] - r["submitted"] for r in res]) * 1e3
    wait = np.array([r["launched"] - r["submitted"] for r in res]) * 1e3
    pts = float(sum(r["points"] for r in res))
    span = max(r["completed"] for r in res) - min(r["submitted"] for r in res)
    offered = rate * mean_pts
    k = max(1, n // 5)
    fifths = [float(np.percentile(wait[i * k:(i + 1) * k], 50)) for i in range(5) if wait[i * k:(i + 1) * k].size]
    return {"rate_req_s": rate, "requests": n, "failed": bad,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
            "queue_wait_p50_ms": float(np.percentile(wait, 50)),
            "queue_wait_p50_first_fifth_ms": float(np.percentile(wait[:k], 50)),
            "queue_wait_p50_last_fifth_ms": float(np.percentile(wait[-k:], 50)),
            "queue_wait_p50_by_fifth_ms": fifths,
            "throughput": pts / span, "offered_points_per_s": offered,
            "utilisation": offered / capacity}


def bench_serving(ctx, args, bj, capacity):
    """Poisson request streams against the continuous-batching server over the
    headline's 64 resident jobs (their grids gridded by the batch): an
    overload at 1.5x the specified rate (the server's own capacity, its
    throughput when it is never idle), the specified 200 req/s, and a declared
    rate at 60 % of the measured batched capacity.  The arena is what a 200 GB
    budget leaves beside the resident data."""
    import tcv
    bj.zero_grids()
    bj.run_batched(graph=True)
    ctx.sync()
    free, total = ctx.mem_info()
    resident = total - free
    budget = 200 * 10**9
    arena = int(min(budget - resident, free - 16 * 2**30))
    specs = bj.specs
    mean_pts = float(np.mean([sp["nVis"] * sp["sSize"] ** 2 for sp in specs]))
    t = bj.tuning
    srv = tcv.Server(ctx, arena, max_inflight=args.serve_inflight, auto_release=True,
                     max_batch_points=int(args.serve_batch_points), tile_w=t.get("tile_w", 0),
                     tile_h=t.get("tile_h", 0), item_entries=t.get("item_entries", 0),
                     grid_cols=t.get("grid_cols", 0), exclusive_points=int(args.serve_exclusive_points))
    out = {"workload": "Poisson requests over the 64 uneven jobs, 50% grid into new grids, 50% degrid "
                       "resident grids, HBM arena admission, overlapping batches",
           "arena_bytes": arena, "resident_bytes": resident, "budget_bytes": budget,
           "max_inflight": args.serve_inflight, "max_batch_points": args.serve_batch_points,
           "exclusive_points": args.serve_exclusive_points,
           "tile": [t.get("tile_w"), t.get("tile_h")], "item_entries": t.get("item_entries", 0),
           "grid_cols": t.get("grid_cols", 0),
           "capacity_points_per_s": capacity, "mean_points_per_request": mean_pts}
    sat = serving_stream(srv, bj, mean_pts, capacity, 300.0, max(300, (args.serve_requests * 9) // 20), 5)
    out["server_capacity_points_per_s"] = sat["throughput"]
    out["server_capacity_over_200_per_s_load"] = sat["throughput"] / (200.0 * mean_pts)
    out["saturated"] = sat
    log(f"serving capacity (300/s offered): {sat['throughput'] / 1e9:.0f} G points/s, "
        f"{out['server_capacity_over_200_per_s_load']:.2f}x the 200/s load")
    out["specified"] = serving_stream(srv, bj, mean_pts, capacity, 200.0, args.serve_requests, 1)
    s = out["specified"]
    log(f"serving @200/s: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms (x{s['p99_over_p50']:.2f}), "
        f"utilisation {s['utilisation']:.2f}, queue wait p50 by fifth "
        + "/".join(f"{x:.1f}" for x in s["queue_wait_p50_by_fifth_ms"]) + " ms")
    rate = 0.6 * capacity / mean_pts
    out["declared_60pct"] = serving_stream(srv, bj, mean_pts, capacity, rate,
                                           max(200, args.serve_requests // 2), 2)
    s = out["declared_60pct"]
    log(f"serving @{rate:.0f}/s: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms "
        f"(x{s['p99_over_p50']:.2f})")
    srv.close()
    return out


# ---- CPU baseline --------------------------------------------------------------------

REF_TIMER = os.path.join(REPO, "oracle", "_ref", "ref_time_attic")


def host_cpus():
    """Logical CPUs of the machine, those this process may run on, and the
    physical cores (distinct (package, core) pairs in /proc/cpuinfo)."""
    phys = set()
    pkg = core = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("physical id"):
                pkg = ln.split(":")[1].strip()
            elif ln.startswith("core id"):
                core = ln.split(":")[1].strip()
            elif not ln.strip():
                if core is not None:
                    phys.add((pkg, core))
                pkg = core = None
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"os_cpu_count": os.cpu_count(), "allowed": allowed, "physical_cores": len(phys) or None}


def cpu_job_rate(spec, C, iu, iv, co, threads, per_thread):
    """tConvolveMT methodology on one job: `threads` instances, each gridding
    then degridding its own `per_thread` visibilities (taken in turn from the
    job's, wrapping) into its own grid, C shared read-only.  Returns (seconds
    to grid, seconds to degrid, kind, what)."""
    import subprocess
    import tempfile
    S = int(spec["sSize"])
    if os.access(REF_TIMER, os.X_OK):
        with tempfile.TemporaryDirectory() as d:
            np.asarray(iu, np.int32).tofile(os.path.join(d, "iu.i32"))
            np.asarray(iv, np.int32).tofile(os.path.join(d, "iv.i32"))
            np.asarray(co, np.int32).tofile(os.path.join(d, "cOffset.i32"))
            np.asarray(spec["data"], np.complex64).tofile(os.path.join(d, "data.c64"))
            np.asarray(C, np.complex64).tofile(os.path.join(d, "C.c64"))
            with open(os.path.join(d, "params.txt"), "w") as f:
                f.write(f"{spec['gSize']} {S // 2}\n")
            run = subprocess.run([REF_TIMER, d, str(threads), str(per_thread)], capture_output=True,
                                 text=True, timeout=600, check=True)
        r = json.loads(run.stdout.strip().splitlines()[-1])
        return (r["grid_s"], r["degrid_s"], "reference",
                "the reference's own gridKernel/degridKernel (attic/tConvolveMPI/Benchmark.cc, "
                "compiled with its flags)")
    import oracle as O
    tg, td = O.cpu_baseline(threads, per_thread, spec["data"], iu, iv, co, C, spec["gSize"], S)
    return tg, td, "port", "the oracle's restatement of gridKernel/degridKernel"


def cpu_baseline(bj, seconds_per_phase=2.5):
    """The CPU path beside the GPU: on a bounded sample of the headline jobs
    (the most-points job of each grid size), tConvolveMPI's reporting
    (attic/tConvolveMPI/tConvolveMPI.cc:82-122): a rate per process (one
    instance per thread here, as tConvolveMT runs them; MPI is absent) and the
    total, on BOX_CPU_SHARE threads, plus one instance alone (1 core)."""
    cpus = host_cpus()
    threads = max(1, min(BOX_CPU_SHARE, cpus["allowed"]))
    picks = {}
    for j, job in enumerate(bj.jobs):
        sp = job["spec"]
        p = sp["nVis"] * sp["sSize"] ** 2
        if sp["gSize"] not in picks or p > picks[sp["gSize"]][1]:
            picks[sp["gSize"]] = (j, p)
    tot = {"pts_n": 0.0, "t_n": 0.0, "pts_1": 0.0, "t_1": 0.0}
    parts = []
    kind = what = None
    for g in sorted(picks):
        job = bj.jobs[picks[g][0]]
        sp = job["spec"]
        S = int(sp["sSize"])
        iu, iv, co = job["vis"].iu.download(), job["vis"].iv.download(), job["vis"].cOffset.download()
        C = job["C"].download()
        # about 1 ns a point per thread for each phase (measured, compiled reference)
        per = int(max(1000, seconds_per_phase / (1.0e-9 * S * S)))
        tg, td, kind, what = cpu_job_rate(sp, C, iu, iv, co, threads, per)
        per1 = max(500, per // 4)
        tg1, td1, _, _ = cpu_job_rate(sp, C, iu, iv, co, 1, per1)
        pts, pts1 = threads * per * S * S, per1 * S * S
        tot["pts_n"] += 2 * pts
        tot["t_n"] += tg + td
        tot["pts_1"] += 2 * pts1
        tot["t_1"] += tg1 + td1
        parts.append({"job": picks[g][0], "gSize": g, "sSize": S, "visibilities_per_instance": per,
                      "grid_s": tg, "degrid_s": td, "one_core_visibilities": per1,
                      "one_core_grid_s": tg1, "one_core_degrid_s": td1})
        del C
    total = tot["pts_n"] / tot["t_n"]
    return {
        "value": total, "unit": "grid points/s", "cores": threads, "kind": kind,
        "per_process": total / threads, "one_core": tot["pts_1"] / tot["t_1"],
        "host": cpus,
        "cores_note": (f"{threads} threads: the GPU box grants one GPU's job {BOX_CPU_SHARE} host "
                       "CPUs whatever the machine's count (host fields)"),
        "sample": (f"the most-points job of each grid size of the 64-job mix; per job {threads} "
                   "instances, each gridding then degridding its own visibilities into its own grid "
                   f"(tConvolveMT methodology, C shared read-only), {what}; and one instance alone"),
        "jobs": parts,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nsamples", type=int, default=1 << 20)
    ap.add_argument("--nchan", type=int, default=8)
    ap.add_argument("--spectral-steps", type=int, default=10)
    ap.add_argument("--tile-w", type=int, default=0)
    ap.add_argument("--tile-h", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--slice-keys", type=int, default=0)
    ap.add_argument("--phases", type=int, default=0)
    ap.add_argument("--xcd-split", type=int, default=-1)
    ap.add_argument("--kernel", type=int, default=0,
                    help="spectral gridder: 1 register tile, 2 LDS tile, 3 register tile with tap pairs")
    ap.add_argument("--grid-phases", type=int, default=0)
    ap.add_argument("--grid-cols", type=int, default=0,
                    help="spectral pair gridder: tile columns per load instruction (0: automatic)")
    ap.add_argument("--degrid-kernel", type=int, default=0,
                    help="spectral degridder: 1 register tile, 2 LDS tile, 3 LDS tile with tap pairs")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--batch-tuning", default="",
                    help="JSON tuning for the batch instead of its autotuner")
    ap.add_argument("--config1-reps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the headline batch (and serving and the CPU baseline, which use it)")
    ap.add_argument("--no-spectral", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-half", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--serve-batch-points", type=float, default=1e10)
    ap.add_argument("--serve-exclusive-points", type=float, default=1e10,
                    help="requests of at least this many grid points run alone (0: never)")
    ap.add_argument("--no-major", action="store_true")
    ap.add_argument("--no-current", action="store_true")
    ap.add_argument("--major-samples", type=int, default=160000)
    ap.add_argument("--major-cycles", type=int, default=5)
    ap.add_argument("--major-minor", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the gridding path here runs on a single "
              f"MI355X (requested --gpus {args.gpus}); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import tcv

    ctx = tcv.Context(0)
    calibration_copy(ctx)
    ceilings = read_ceilings(ctx)
    args._ceilings = ceilings
    log("read ceilings GB/s: " + ", ".join(f"{k} {v:.0f}" for k, v in ceilings.items()))

    if args.no_batch:                  # profiling and tuning runs of the other sections
        result = {"metric": "grid points/s (grid+degrid) batched vs plain, 1 GPU; p50/p99 job latency",
                  "value": None, "note": "headline batch not run (--no-batch)"}
        args.no_serve = args.no_cpu = True
        bj = None
    else:
        bj, batch = bench_batch(ctx, args, ceilings)
    result = result if args.no_batch else {
        "metric": "grid points/s (grid+degrid) batched vs plain, 1 GPU; p50/p99 job latency",
        "value": batch["value"], "unit": "grid points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": batch["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (complex64)",
        "data": "synthetic (seeded: numpy PCG64 per-job u,v,w uniform [-B,B], complex Gaussian data)",
        "config": {"workload": batch["workload"], "jobs": batch["jobs"],
                   "visibilities": batch["visibilities"],
                   "grid_points_per_pass": batch["grid_points_per_pass"],
                   "kernel_widths": batch["kernel_widths"], "tuning": batch["tuning"]},
        "speedup_batched_vs_plain": batch["speedup_batched_vs_plain"],
        "plain_per_job_value": batch["plain_per_job"]["value"],
        "roofline": batch["roofline"],
        "batch": batch,
    }
    if not args.no_serve:
        result["serving"] = bench_serving(ctx, args, bj, batch["value"])
    if not args.no_cpu:
        cb = cpu_baseline(bj)
        result["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value'] / 1e6:.0f} M points/s total on {cb['cores']} threads, "
            f"{cb['per_process'] / 1e6:.0f} per process, {cb['one_core'] / 1e6:.0f} on one core")
    del bj
    ctx.sync()
    tuning = dict(tile_w=args.tile_w, tile_h=args.tile_h, chunk=args.chunk,
                  slice_keys=args.slice_keys, phases=args.phases, xcd_split=args.xcd_split,
                  kernel=args.kernel, degrid_kernel=args.degrid_kernel, grid_phases=args.grid_phases,
                  grid_cols=args.grid_cols)
    if not args.no_config1:
        c1 = {v: bench_config1(ctx, v, args.config1_reps, tuning) for v in ("reference", "wide")}
        result["config1"] = c1
        r = c1["reference"]["fp32"]
        result["config1_grid_Mpts"] = r["grid_Mpts"]
        result["config1_degrid_Mpts"] = r["degrid_Mpts"]
        ra = c1["reference"]["fp32_autotuned"]
        result["config1_autotuned_grid_Mpts"] = ra["grid_Mpts"]
        result["config1_autotuned_degrid_Mpts"] = ra["degrid_Mpts"]
        result["p100_grid_Mpts"] = P100_GRID_MPTS
        result["p100_degrid_Mpts"] = P100_DEGRID_MPTS
    if not args.no_spectral:
        result["spectral"] = bench_spectral(ctx, args, tuning, args.spectral_steps)
    # the current generation's runs sit with config 1 (the other per-process
    # comparison): --no-config1 leaves both out
    if not args.no_config1 and not args.no_current:
        result["current"] = bench_current(ctx, args.config1_reps)
    if not args.no_major:
        result["major_cycle"] = bench_major(ctx, args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
