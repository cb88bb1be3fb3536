"""Benchmark: W-projection gridding + degridding throughput on one MI355X.

Workload (spectral scale): 1 048 576 samples x 8 channels, u, v, w uniform in
[-6000, 6000] m, complex Gaussian data, 4096^2 complex64 grid, wSize 129,
overSample 8, cellSize 15 wavelengths (support 64, 129 x 129 taps).

One step = re-bin the visibilities on the device (plan build) + zero the grid
+ grid every visibility + degrid every visibility, all on device-resident
inputs.  value = (grid points gridded + grid points degridded) / step time.

Prints one JSON line.  Extra fields: the plain one-workgroup-per-visibility
kernels timed on the same data, the roofline of the dominant kernel and the
CPU baseline on the host cores, on a bounded sample: the reference's own
gridKernel/degridKernel compiled from its sources when that checker binary was
built, else the oracle's restatement of them.
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

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def alg_bytes(points, nVis, gSize, op):
    """Algorithmic HBM bytes: one complex64 C tap per grid point, the grid read
    (degrid) or read+written (grid) once, 20 B per visibility (iu, iv, cOffset,
    data)."""
    grid_bytes = (16 if op == "grid" else 8) * gSize * gSize
    return 8 * points + grid_bytes + 20 * nVis


TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_grid_degrid_tiled_spectral.md")
TUNING_KEYS = ("tile_w", "tile_h", "chunk", "slice_keys", "phases", "xcd_split", "kernel",
               "degrid_kernel", "grid_phases")


def measured_traffic(kernel, info, field="bytes_per_launch"):
    """Bytes per launch of `kernel` from the committed counter summary, when it
    was measured with the same tuning; else None.  bytes_per_launch: between
    L2 and the memory fabric (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration); l2_read_bytes_per_launch: from L2 into the L1s
    (TCP_TCC_READ_REQ times the request size the copy kernel shows)."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    rec = {}
    for ln in lines:
        if ln.startswith("traffic-record:"):
            rec = json.loads(ln.split(":", 1)[1])
    if not rec or any(rec.get("tuning", {}).get(k) != info.get(k) for k in TUNING_KEYS):
        return None
    return rec.get(field, {}).get(kernel)


def read_ceilings(ctx):
    """Streaming read bandwidth (GB/s) of the device's levels, measured live:
    a 2 MiB buffer (held by the L2s), 64 MiB (Infinity Cache) and 1 GiB (HBM)."""
    import numpy as np
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
    import numpy as np
    a = ctx.empty(nbytes // 4, np.uint32)
    b = ctx.empty(nbytes // 4, np.uint32)
    a.zero()
    for _ in range(reps):
        b.copy_from(a)
    ctx.sync()
    a.free()
    b.free()


REF_TIMER = os.path.join(REPO, "oracle", "_ref", "ref_time_attic")


def cpu_baseline(prob, iu, iv, co, seconds_target=16.0):
    """tConvolveMT-style CPU rate on the box's host cores: one instance per
    thread, each gridding then degridding its own share of the workload's
    visibilities into its own grid, C shared read-only.  The reference's own
    gridKernel/degridKernel compiled from its sources (oracle/_ref, kind
    "reference") when that binary was built, else the oracle's restatement
    (kind "port")."""
    import subprocess
    import tempfile
    cores = min(os.cpu_count() or 1, 16)
    S = int(prob.sSize[0])
    # measured with 16 threads: about 0.8 ns per point and thread to grid and
    # 1.0 ns to degrid (compiled reference), so a share takes ~2 ns a point
    per_thread = int(max(2000, seconds_target / (2.0e-9 * S * S)))
    per_thread = min(per_thread, iu.size)
    kind = "port"
    if os.access(REF_TIMER, os.X_OK):
        with tempfile.TemporaryDirectory() as d:
            np.asarray(iu, np.int32).tofile(os.path.join(d, "iu.i32"))
            np.asarray(iv, np.int32).tofile(os.path.join(d, "iv.i32"))
            np.asarray(co, np.int32).tofile(os.path.join(d, "cOffset.i32"))
            np.asarray(prob.data, np.complex64).tofile(os.path.join(d, "data.c64"))
            np.asarray(prob.C, np.complex64).tofile(os.path.join(d, "C.c64"))
            with open(os.path.join(d, "params.txt"), "w") as f:
                f.write(f"{prob.gSize} {S // 2}\n")
            run = subprocess.run([REF_TIMER, d, str(cores), str(per_thread)], capture_output=True,
                                 text=True, timeout=600, check=True)
        r = json.loads(run.stdout.strip().splitlines()[-1])
        tg, td = r["grid_s"], r["degrid_s"]
        kind = "reference"
        what = ("the reference's own gridKernel/degridKernel (attic/tConvolveMPI/Benchmark.cc, "
                "compiled with its flags)")
    else:
        import oracle as O
        tg, td = O.cpu_baseline(cores, per_thread, prob.data, iu, iv, co, prob.C, prob.gSize, S)
        what = "the oracle's restatement of gridKernel/degridKernel"
    pts = cores * per_thread * S * S
    rate = 2 * pts / (tg + td)
    return {
        "value": rate, "unit": "grid points/s", "cores": cores, "kind": kind,
        "sample": (f"{cores} threads x {per_thread} visibilities of the same workload, each "
                   f"thread gridding into its own {prob.gSize}^2 grid then degridding "
                   f"(tConvolveMT methodology, C shared read-only), {what}; grid {tg:.2f} s, "
                   f"degrid {td:.2f} s"),
        "grid_rate": pts / tg, "degrid_rate": pts / td,
    }


def bench_batch(ctx, args):
    """The uneven 64-job batch: the same jobs run as one persistent
    hipGraph-captured launch (batched), as one tiled launch per job, and with
    the plain one-workgroup-per-visibility kernels, one launch per job.  Each
    mode re-bins (where it bins at all), grids and degrids every job."""
    from tcv import setup as S
    from tcv.gridder import BatchJobs
    t0 = time.time()
    specs = S.uneven_batch(nJobs=args.batch_jobs)
    bj = BatchJobs(ctx, specs)
    log(f"batch: {len(specs)} jobs, {sum(s['nVis'] for s in specs)} visibilities, "
        f"{bj.points / 1e9:.1f} G points per pass, set up in {time.time() - t0:.1f} s")
    pts = 2 * bj.points

    def timed(fn, reps):
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            bj.zero_grids()
            fn()
        return ctx.timer_stop() / reps

    timed(lambda: bj.run_batched(graph=True), 1)          # capture + warm
    t_batch = timed(lambda: bj.run_batched(graph=True), args.batch_steps)
    timed(lambda: bj.run_tiled(), 1)
    t_tiled = timed(lambda: bj.run_tiled(), args.batch_steps)
    t_plain = timed(lambda: bj.run_plain(), 1)
    out = {
        "workload": "64 jobs: visibilities log-uniform [1e4, 2e6], grids {1024,2048,4096}^2, "
                    "wSize {33,65,129}, baselines 1000-6000 m, complex64, tightest-fit cellSize",
        "jobs": len(specs), "visibilities": int(sum(s["nVis"] for s in specs)),
        "points_per_step": pts,
        "batched": {"ms": t_batch, "value": pts / (t_batch * 1e-3)},
        "tiled_per_job": {"ms": t_tiled, "value": pts / (t_tiled * 1e-3)},
        "plain_per_job": {"ms": t_plain, "value": pts / (t_plain * 1e-3)},
    }
    out["speedup_batched_vs_plain"] = t_plain / t_batch
    out["speedup_batched_vs_tiled"] = t_tiled / t_batch
    log(f"batch: batched {t_batch:.1f} ms, tiled/job {t_tiled:.1f} ms, plain/job {t_plain:.1f} ms "
        f"-> batched/plain {out['speedup_batched_vs_plain']:.2f}x")
    del bj
    return out


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
    log(f"major: {out['per_cycle_ms']:.1f} ms per cycle; stages "
        + ", ".join(f"{k} {v:.1f}" for k, v in out["stages_ms"].items()))
    del mc
    return out


def serving_stream(srv, bj, mean_pts, capacity, rate, n, seed):
    """One Poisson stream of n requests at `rate` per second against `srv`:
    requests drawn from the resident jobs, 50% grid into a new grid, 50%
    degrid the job's resident grid.  Latency = submission to completion."""
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
        srv.release(r)
    bad = sum(r["status"] != 0 for r in res)
    lat = np.array([r["completed"] - r["submitted"] for r in res]) * 1e3
    wait = np.array([r["launched"] - r["submitted"] for r in res]) * 1e3
    pts = float(sum(r["points"] for r in res))
    span = max(r["completed"] for r in res) - min(r["submitted"] for r in res)
    offered = rate * mean_pts
    return {"rate_req_s": rate, "requests": n, "failed": bad,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
            "queue_wait_p50_ms": float(np.percentile(wait, 50)),
            "throughput": pts / span, "offered_points_per_s": offered,
            "utilisation": offered / capacity}


def serving_setup(ctx):
    """The 64 uneven jobs resident on the device (grids gridded once for the
    degrid requests) and the arena that keeps resident data plus requests
    within a 200 GB budget."""
    from tcv import setup as S
    from tcv.gridder import BatchJobs
    specs = S.uneven_batch(nJobs=64)
    bj = BatchJobs(ctx, specs)
    bj.zero_grids()
    bj.run_batched(graph=False)
    ctx.sync()
    free, total = ctx.mem_info()
    resident = total - free
    budget = 200 * 10**9
    arena = int(min(budget - resident, free - 16 * 2**30))
    mean_pts = float(np.mean([sp["nVis"] * sp["sSize"] ** 2 for sp in specs]))
    return bj, arena, resident, budget, mean_pts, specs


def bench_serving(ctx, args, capacity):
    """Poisson request streams against the continuous-batching server: the
    specified rate, and a declared rate at 60% of the measured batched
    capacity.  Requests follow the uneven 64-job size distribution."""
    import tcv
    from tcv.gridder import batch_tile_w
    bj, arena, resident, budget, mean_pts, specs = serving_setup(ctx)
    tile_w = batch_tile_w([sp["sSize"] for sp in specs])
    srv = tcv.Server(ctx, arena, max_inflight=args.serve_inflight, auto_release=True,
                     max_batch_points=int(args.serve_batch_points), tile_w=tile_w, tile_h=64)
    out = {"workload": "Poisson requests over the 64 uneven jobs, 50% grid into new grids, 50% degrid "
                       "resident grids, HBM arena admission, overlapping batches",
           "arena_bytes": arena, "resident_bytes": resident, "budget_bytes": budget,
           "max_inflight": args.serve_inflight, "max_batch_points": args.serve_batch_points,
           "tile": [tile_w, 64],
           "capacity_points_per_s": capacity, "mean_points_per_request": mean_pts}
    out["specified"] = serving_stream(srv, bj, mean_pts, capacity, 200.0, args.serve_requests, 1)
    log(f"serving @200/s: p50 {out['specified']['p50_ms']:.1f} ms p99 {out['specified']['p99_ms']:.1f} ms, "
        f"utilisation {out['specified']['utilisation']:.2f}")
    rate = 0.6 * capacity / mean_pts
    out["declared_60pct"] = serving_stream(srv, bj, mean_pts, capacity, rate,
                                           max(200, args.serve_requests // 2), 2)
    log(f"serving @{rate:.0f}/s: p50 {out['declared_60pct']['p50_ms']:.1f} ms "
        f"p99 {out['declared_60pct']['p99_ms']:.1f} ms")
    srv.close()
    del bj
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nsamples", type=int, default=1 << 20)
    ap.add_argument("--nchan", type=int, default=8)
    ap.add_argument("--tile-w", type=int, default=0)
    ap.add_argument("--tile-h", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--slice-keys", type=int, default=0)
    ap.add_argument("--phases", type=int, default=0)
    ap.add_argument("--xcd-split", type=int, default=-1)
    ap.add_argument("--kernel", type=int, default=0,
                    help="gridder: 1 register tile, 2 LDS tile, 3 register tile with tap pairs; 0 auto")
    ap.add_argument("--grid-phases", type=int, default=0,
                    help="C-slice phases of the gridder's work items (0 auto)")
    ap.add_argument("--degrid-kernel", type=int, default=0,
                    help="degridder: 1 register tile, 2 LDS tile, 3 LDS tile with tap pairs; 0 auto")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-half", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--serve-batch-points", type=float, default=1e10)
    ap.add_argument("--no-major", action="store_true")
    ap.add_argument("--major-samples", type=int, default=160000)
    ap.add_argument("--major-cycles", type=int, default=5)
    ap.add_argument("--major-minor", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the gridding path here runs on a single "
              f"MI355X (requested --gpus {args.gpus}); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import tcv
    from tcv import setup as S

    t0 = time.time()
    prob = S.spectral_problem(nSamples=args.nsamples, nChan=args.nchan)
    log(f"problem built in {time.time() - t0:.1f} s: {prob.nVis} visibilities, "
        f"C {prob.C.nbytes / 2**30:.2f} GiB")
    ctx = tcv.Context(0)
    tuning = dict(tile_w=args.tile_w, tile_h=args.tile_h, chunk=args.chunk,
                  slice_keys=args.slice_keys, phases=args.phases, xcd_split=args.xcd_split,
                  kernel=args.kernel, degrid_kernel=args.degrid_kernel, grid_phases=args.grid_phases)
    job = tcv.GridJob(ctx, prob, tuning=tuning)
    assert job.nBad == 0, f"{job.nBad} visibilities out of range"
    points = job.points
    info = job.plan.info()
    log(f"plan: {info}")

    def step():
        job.plan.build()
        job.zero_grid()
        job.grid("tiled")
        job.degrid("tiled")

    calibration_copy(ctx)
    ceilings = read_ceilings(ctx)
    log("read ceilings GB/s: " + ", ".join(f"{k} {v:.0f}" for k, v in ceilings.items()))
    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.set_profiling(True)
    ctx.reset_stats()
    ctx.sync()
    w0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        step()
    ms_total = ctx.timer_stop()
    wall = time.perf_counter() - w0
    stats = {f: ctx.kernel_stats(f) for f in ("plan", "grid_tiled", "degrid_tiled")}
    ctx.set_profiling(False)
    ms_step = ms_total / args.steps
    value = 2 * points / (ms_step * 1e-3)
    log(f"tiled: {ms_step:.2f} ms/step, {value / 1e9:.1f} G points/s (wall {wall:.2f} s)")

    # roofline of the dominant kernel
    fams = {k: v for k, v in stats.items() if k != "plan" and v[0] > 0}
    dom = max(fams, key=lambda k: fams[k][1])
    n_l, ms_l = fams[dom]
    avg_ms = ms_l / n_l
    op = "grid" if dom.startswith("grid") else "degrid"
    bytes_l = alg_bytes(points, prob.nVis, prob.gSize, op)
    achieved = bytes_l / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom, info)
    # the reference's own streaming model: every grid point moves its C tap and
    # its grid pixel (read + write when gridding, read when degridding)
    ref_model = (24 if op == "grid" else 16) * points
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": dom,
                "avg_launch_ms": avg_ms, "alg_bytes_per_launch": bytes_l,
                "measured_hbm_GBps": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                "reference_model_GBps": ref_model / (avg_ms * 1e-3) / 1e9,
                "read_ceilings_GBps": ceilings}
    l2_bytes = measured_traffic(dom, info, "l2_read_bytes_per_launch")
    if l2_bytes:
        roofline["l2_read_bytes_per_launch"] = l2_bytes
        roofline["l2_read_GBps"] = l2_bytes / (avg_ms * 1e-3) / 1e9
        roofline["l2_read_frac_of_l2_ceiling"] = roofline["l2_read_GBps"] / ceilings["l2"]
    per_kernel = {k: {"launches": v[0], "avg_ms": (v[1] / v[0]) if v[0] else None}
                  for k, v in stats.items()}

    result = {
        "metric": "grid points/s (grid+degrid) batched vs plain, 1 GPU; p50/p99 job latency",
        "value": value, "unit": "grid points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (complex64)",
        "data": "synthetic (seeded: numpy PCG64 u,v,w uniform [-6000,6000] m, complex Gaussian data)",
        "config": {"workload": "spectral: 1048576 samples x 8 channels, 4096^2 grid, wSize 129, "
                               "overSample 8, support 64 (129x129 taps), tiled gridder+degridder",
                   "nVis": prob.nVis, "gSize": prob.gSize, "sSize": int(prob.sSize[0]),
                   "grid_points_per_pass": points, "tuning": info},
        "roofline": roofline,
        "kernels": per_kernel,
    }

    if not args.no_half:
        # FP16 storage of C and the visibilities (per-w-plane scales, FP32
        # accumulation): half the bytes per grid point, on a plan whose automatic
        # choices (tile shape, phases) are made for 4-byte taps
        Ch = tcv.HalfTable(ctx, job.C, prob.wSize, int(prob.sSize[0]) ** 2 * 64)
        dh = tcv.HalfVis(ctx, job.data)
        gh = ctx.zeros((prob.gSize, prob.gSize), np.complex64)
        oh = ctx.empty(prob.nVis, np.complex64)
        plan16 = tcv.Plan(ctx, job.vis, **dict(tuning, tap_bytes=4))
        plan16.grid_half(dh, Ch, gh)
        plan16.degrid_half(gh, Ch, oh)
        reps = max(1, min(args.steps, 3))
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            gh.zero()
            plan16.grid_half(dh, Ch, gh)
        t_gh = ctx.timer_stop() / reps
        ctx.timer_start()
        for _ in range(reps):
            plan16.degrid_half(gh, Ch, oh)
        t_dh = ctx.timer_stop() / reps
        # agreement with the FP32 result of the timed steps
        g32 = job.grid_d.download()
        g16 = gh.download()
        err = float(np.abs(g16 - g32).max() / np.abs(g32).max())
        info16 = plan16.info()
        result["half"] = {"value": 2 * points / ((t_gh + t_dh) * 1e-3), "grid_ms": t_gh,
                          "degrid_ms": t_dh, "alg_bytes_per_point": 4,
                          "grid_max_rel_diff_vs_fp32": err,
                          "tuning": {k: info16[k] for k in TUNING_KEYS + ("tap_bytes",)}}
        log(f"fp16 storage: grid {t_gh:.1f} ms, degrid {t_dh:.1f} ms, "
            f"{result['half']['value'] / 1e9:.1f} G points/s, max |fp16 - fp32| / max {err:.1e}")
        plan16.close()
        del Ch, dh, gh, oh

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
        result["plain"] = {"value": plain, "grid_ms": t_gp, "degrid_ms": t_dp,
                           "speedup_tiled_vs_plain": value / plain}
        log(f"plain: grid {t_gp:.1f} ms, degrid {t_dp:.1f} ms, {plain / 1e9:.1f} G points/s")

    if not args.no_batch:
        del job
        ctx.sync()
        result["batch"] = bench_batch(ctx, args)
        if not args.no_serve:
            result["serving"] = bench_serving(ctx, args, result["batch"]["batched"]["value"])

    if not args.no_major:
        result["major_cycle"] = bench_major(ctx, args)

    if not args.no_cpu:
        import oracle as O
        iu, iv, co = O.attic_initCOffset(prob.u, prob.v, prob.w, prob.freq, prob.cellSize,
                                         prob.wCellSize, prob.wSize, prob.gSize, prob.support, 8)
        cb = cpu_baseline(prob, iu, iv, co)
        result["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value'] / 1e6:.0f} M points/s on {cb['cores']} cores")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
