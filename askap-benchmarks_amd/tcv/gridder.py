"""Device-resident gridding jobs: offsets on the device, then plain or tiled
grid / degrid passes.  Everything here calls libtcv; nothing runs on the CPU
except argument marshalling."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import lib
from .device import Context, DeviceArray, Plan, TcvError, Visibilities, degrid_plain, grid_plain
from .setup import Problem

P = ct.c_void_p


def init_offsets(ctx: Context, u: DeviceArray, v: DeviceArray, w: DeviceArray, freq,
                 cellSize: float, wCellSize: float, wSize: int, gSize: int, overSample: int,
                 sSize, cOffset0=None, want_wplane: bool = True, allow_out_of_range=False):
    """Benchmark::initCOffset on the device (FP64, bit-exact).

    attic/tConvolveMPI/Benchmark.cc:281-324 when cOffset0 is None, else
    current/tConvolveMPI/Benchmark.cc:597-681.  Returns device arrays
    (iu, iv, wPlane or None, cOffset), numPerPlane and the out-of-range count."""
    freq = np.ascontiguousarray(freq, np.float64)
    sSize = np.ascontiguousarray(np.atleast_1d(sSize), np.int32)
    co0 = None if cOffset0 is None else np.ascontiguousarray(cOffset0, np.int32)
    nSamples = u.size
    n = nSamples * freq.size
    iu = ctx.empty(n, np.int32)
    iv = ctx.empty(n, np.int32)
    wp = ctx.empty(n, np.int32) if want_wplane else None
    co = ctx.empty(n, np.int32)
    npp = np.zeros(wSize, np.int64)
    bad = ct.c_longlong()
    st = lib().tcv_init_offsets(ctx.h, u.ptr, v.ptr, w.ptr, nSamples, freq.ctypes.data_as(P),
                                freq.size, cellSize, wCellSize, wSize, gSize, overSample,
                                sSize.ctypes.data_as(P),
                                None if co0 is None else co0.ctypes.data_as(P),
                                iu.ptr, iv.ptr, None if wp is None else wp.ptr, co.ptr,
                                npp.ctypes.data_as(P), ct.byref(bad))
    if st and not (st == 2 and allow_out_of_range):
        ctx.check(st, "init_offsets")
    return iu, iv, wp, co, npp, bad.value


class GridJob:
    """One gridding/degridding job resident on the device.

    Holds the visibility geometry (offsets computed on the device from u, v,
    w), the C(w) table, the data, the grid and the degridded output."""

    def __init__(self, ctx: Context, prob: Problem, tuning: dict | None = None,
                 make_plan: bool = True):
        self.ctx = ctx
        self.prob = prob
        self.d_u = ctx.to_device(prob.u)
        self.d_v = ctx.to_device(prob.v)
        self.d_w = ctx.to_device(prob.w)
        self.iu, self.iv, self.wp, self.co, self.numPerPlane, self.nBad = init_offsets(
            ctx, self.d_u, self.d_v, self.d_w, prob.freq, prob.cellSize, prob.wCellSize,
            prob.wSize, prob.gSize, prob.overSample, prob.sSize, prob.cOffset0,
            want_wplane=prob.cOffset0 is not None)
        self.C = ctx.to_device(prob.C)
        self.data = ctx.to_device(prob.data)
        self.grid_d = ctx.zeros((prob.gSize, prob.gSize), np.complex64)
        self.out = ctx.zeros(prob.nVis, np.complex64)
        self.vis = Visibilities(ctx, self.iu, self.iv, self.co, prob.sSize, prob.gSize,
                                prob.C.size, wPlane=self.wp)
        self.plan = Plan(ctx, self.vis, **(tuning or {})) if make_plan else None

    @property
    def points(self) -> int:
        s = self.prob.sSize.astype(np.int64)
        if self.prob.cOffset0 is None:
            return int(self.prob.nVis) * int(s[0]) ** 2
        return int((self.numPerPlane * s * s).sum())

    def zero_grid(self):
        self.grid_d.zero()

    def grid(self, method: str = "tiled"):
        if method == "tiled":
            self.plan.grid(self.data, self.C, self.grid_d)
        elif method == "plain":
            grid_plain(self.ctx, self.vis, self.data, self.C, self.grid_d)
        else:
            raise ValueError(method)

    def degrid(self, method: str = "tiled"):
        if method == "tiled":
            self.plan.degrid(self.grid_d, self.C, self.out)
        elif method == "plain":
            degrid_plain(self.ctx, self.vis, self.grid_d, self.C, self.out)
        else:
            raise ValueError(method)

    def offsets_host(self):
        wp = self.wp.download() if self.wp is not None else None
        return self.iu.download(), self.iv.download(), wp, self.co.download()


__all__ = ["init_offsets", "GridJob", "TcvError", "batch_tile_w"]


WIDE_TILE_MIN_WIDTH = 89   # the library's rule for a single plan (FP32 tables)


def batch_tile_w(widths, tuning: dict | None = None) -> int:
    """The one tile width shared by a batch (or a server's requests): 128 when
    the widest kernel is at least 89 wide (the rule a single job's plan
    follows), every kernel is at least 2 wide and the pair kernels (code 3, or
    automatic) are used; else 64."""
    t = tuning or {}
    pairs = t.get("kernel", 0) in (0, 3) and t.get("degrid_kernel", 0) in (0, 3)
    widths = [int(w) for w in widths]
    return 128 if pairs and max(widths) >= WIDE_TILE_MIN_WIDTH and min(widths) >= 2 else 64


class BatchJobs:
    """A set of independent jobs resident on the device (the uneven batch).

    Each job gets its own grid, C(w) table (generated on the device), data,
    offsets and plan; run_* execute the whole set one way or another so the
    three execution modes can be compared on identical inputs:
      plain    one launch per job of the one-workgroup-per-visibility kernels
      tiled    one tiled (binned) launch per job
      batched  one persistent launch over all jobs' work items, captured in a
               hipGraph together with the re-binning
    """

    def __init__(self, ctx: Context, specs, tuning: dict | None = None):
        from .device import Batch, attic_initC_device
        self.ctx = ctx
        self.specs = specs
        self.jobs = []
        tuning = dict(tuning or {})
        tuning.setdefault("tile_w", batch_tile_w([sp["sSize"] for sp in specs], tuning))
        for sp in specs:
            sup, wcs, C = attic_initC_device(ctx, float(sp["freq"][0]), sp["cellSize"], sp["wSize"],
                                             int(sp["baseline"]))
            assert 2 * sup + 1 == sp["sSize"], (sup, sp["sSize"])
            du, dv, dw = (ctx.to_device(sp[k]) for k in ("u", "v", "w"))
            iu, iv, _, co, npp, bad = init_offsets(ctx, du, dv, dw, sp["freq"], sp["cellSize"], wcs,
                                                   sp["wSize"], sp["gSize"], 8, [sp["sSize"]], None,
                                                   want_wplane=False)
            if bad:
                raise ValueError(f"job {sp['job']}: {bad} visibilities out of range")
            vis = Visibilities(ctx, iu, iv, co, [sp["sSize"]], sp["gSize"], C.size)
            job = dict(spec=sp, C=C, data=ctx.to_device(sp["data"]), vis=vis,
                       grid=ctx.zeros((sp["gSize"], sp["gSize"]), np.complex64),
                       out=ctx.zeros(sp["nVis"], np.complex64), wCellSize=wcs,
                       plan=Plan(ctx, vis, **tuning))
            del du, dv, dw
            self.jobs.append(job)
        self.batch = Batch(ctx, [(j["plan"], j["data"], j["C"], j["grid"], j["out"])
                                 for j in self.jobs])
        self.tuning = tuning
        self.points = int(sum(sp["nVis"] * sp["sSize"] ** 2 for sp in specs))

    def set_tuning(self, tuning: dict | None):
        """Re-plans every job with `tuning` (one tile shape for the whole batch;
        missing fields automatic) and rebuilds the batch over the new plans."""
        from .device import Batch
        tuning = dict(tuning or {})
        tuning.setdefault("tile_w", batch_tile_w([sp["sSize"] for sp in self.specs], tuning))
        self.batch.close()
        for j in self.jobs:
            j["plan"].close()
            j["plan"] = Plan(self.ctx, j["vis"], **tuning)
        self.batch = Batch(self.ctx, [(j["plan"], j["data"], j["C"], j["grid"], j["out"])
                                      for j in self.jobs])
        self.tuning = tuning

    def autotune(self, candidates=None, reps: int = 2):
        """The batch's autotuner: times the captured batched step (re-bin + grid
        + degrid of every job) under each candidate tuning and keeps the
        fastest.  Default candidates: the tile shapes a batch can share (128 x 64
        and 64 x 128 with the pair kernels when every kernel is at least 2 wide,
        else 64 x 64) and, for each, the fewest entries per work item that each
        job's automatic C-slice phase count keeps (finer phases keep more of a
        job's slices in the L2s; each work item zeroes and flushes its tile),
        and on 128 x 64 tiles the pair gridder's columns per load instruction.
        Returns (best tuning, [(tuning, ms per step), ...])."""
        if candidates is None:
            widths = [sp["sSize"] for sp in self.specs]
            if min(widths) >= 2:
                pair = dict(kernel=3, degrid_kernel=3)
                wide = dict(pair, tile_w=128, tile_h=64)
                candidates = ([dict(wide, item_entries=n, grid_cols=32) for n in (64, 40, 32, 24)]
                              + [dict(wide, item_entries=32, grid_cols=c) for c in (64, 128)]
                              + [dict(pair, tile_w=64, tile_h=128, item_entries=32)])
            else:
                candidates = [dict(tile_w=64, tile_h=64, item_entries=n) for n in (64, 32)]
        table = []
        for cand in candidates:
            self.set_tuning(cand)
            self.zero_grids()
            self.run_batched(graph=True)          # capture + warm
            self.ctx.sync()
            self.ctx.timer_start()
            for _ in range(reps):
                self.zero_grids()
                self.run_batched(graph=True)
            table.append((dict(cand), self.ctx.timer_stop() / reps))
        best = min(table, key=lambda t: t[1])[0]
        self.set_tuning(best)
        return best, table

    def zero_grids(self):
        for j in self.jobs:
            j["grid"].zero()

    def run_plain(self):
        for j in self.jobs:
            grid_plain(self.ctx, j["vis"], j["data"], j["C"], j["grid"])
        for j in self.jobs:
            degrid_plain(self.ctx, j["vis"], j["grid"], j["C"], j["out"])

    def run_tiled(self, rebuild: bool = True):
        for j in self.jobs:
            if rebuild:
                j["plan"].build()
            j["plan"].grid(j["data"], j["C"], j["grid"])
            j["plan"].degrid(j["grid"], j["C"], j["out"])

    def run_batched(self, graph: bool = True, rebuild: bool = True):
        ops = (1 if rebuild else 0) | 2 | 4
        if graph:
            self.batch.run_graph(ops)
        else:
            if rebuild:
                self.batch.rebuild()
            self.batch.grid()
            self.batch.degrid()
