"""Device context, device arrays, geometry and plans over the C ABI."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import STATUS, Geometry, Job, Request, Result, ServerConfig, Tuning, lib

P = ct.c_void_p


class TcvError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        self.code = STATUS.get(status, str(status))
        super().__init__(f"{self.code}: {message}")


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


class Context:
    """One device, one stream (tcv_ctx)."""

    def __init__(self, device: int = 0):
        h = P()
        st = lib().tcv_create(device, ct.byref(h))
        if st:
            raise TcvError(st, f"tcv_create(device={device}) failed")
        self.h = h
        self.device = device

    # -- errors -------------------------------------------------------------
    def check(self, st: int, what: str = ""):
        if st:
            msg = lib().tcv_last_error(self.h)
            raise TcvError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            lib().tcv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def cus(self) -> int:
        return lib().tcv_device_cus(self.h)

    def sync(self):
        self.check(lib().tcv_synchronize(self.h), "synchronize")

    # -- memory -------------------------------------------------------------
    def empty(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def zeros(self, shape, dtype) -> "DeviceArray":
        a = DeviceArray(self, shape, dtype)
        a.zero()
        return a

    def to_device(self, host) -> "DeviceArray":
        host = np.ascontiguousarray(host)
        a = DeviceArray(self, host.shape, host.dtype)
        a.upload(host)
        return a

    # -- timing -------------------------------------------------------------
    def timer_start(self):
        self.check(lib().tcv_timer_start(self.h), "timer_start")

    def timer_stop(self) -> float:
        ms = ct.c_double()
        self.check(lib().tcv_timer_stop(self.h, ct.byref(ms)), "timer_stop")
        return ms.value

    def set_profiling(self, on: bool):
        self.check(lib().tcv_set_profiling(self.h, int(on)), "set_profiling")

    def kernel_stats(self, family: str):
        n = ct.c_long()
        ms = ct.c_double()
        self.check(lib().tcv_kernel_stats(self.h, family.encode(), ct.byref(n), ct.byref(ms)),
                   "kernel_stats")
        return n.value, ms.value

    def reset_stats(self):
        self.check(lib().tcv_reset_stats(self.h), "reset_stats")

    def mem_info(self):
        f, t = ct.c_size_t(), ct.c_size_t()
        self.check(lib().tcv_mem_info(self.h, ct.byref(f), ct.byref(t)), "mem_info")
        return f.value, t.value

    def read_bandwidth(self, buf: "DeviceArray", rounds: int = 256) -> float:
        """Streaming read bandwidth (GB/s) over `buf` (power-of-two bytes): the
        L2's when it fits there, HBM's for a large buffer (tcv_read_bandwidth)."""
        g = ct.c_double()
        self.check(lib().tcv_read_bandwidth(self.h, buf.ptr, buf.nbytes, int(rounds), ct.byref(g)),
                   "read_bandwidth")
        return g.value

    def range_errors(self) -> int:
        c = ct.c_longlong()
        self.check(lib().tcv_range_errors(self.h, ct.byref(c)), "range_errors")
        return c.value


class DeviceArray:
    """A device buffer owned by a Context (freed with the object)."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        h = P()
        ctx.check(lib().tcv_malloc(ctx.h, self.nbytes, ct.byref(h)), f"malloc {self.nbytes} B")
        self.ptr = h

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.nbytes != self.nbytes:
            raise ValueError(f"upload of {host.nbytes} B into a {self.nbytes} B buffer")
        self.ctx.check(lib().tcv_upload(self.ctx.h, self.ptr, _p(host), self.nbytes), "upload")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(lib().tcv_download(self.ctx.h, _p(out), self.ptr, self.nbytes), "download")
        return out

    def zero(self):
        self.ctx.check(lib().tcv_zero(self.ctx.h, self.ptr, self.nbytes), "zero")

    def copy_from(self, src: "DeviceArray"):
        """Device-to-device copy by the library's streaming kernel (tcv_copy)."""
        if src.nbytes != self.nbytes:
            raise ValueError("copy_from: size mismatch")
        self.ctx.check(lib().tcv_copy(self.ctx.h, self.ptr, src.ptr, self.nbytes), "copy")

    def free(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            lib().tcv_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Visibilities:
    """Geometry of one job: device index arrays plus the per-plane widths.

    Mirrors the reference's per-visibility offsets (attic `Sample` iu/iv/cOffset,
    attic/tConvolveMPI/Benchmark.h:48-53; current iu/iv/wPlane/cOffset,
    current/tConvolveMPI/Benchmark.h:96-103)."""

    def __init__(self, ctx: Context, iu, iv, cOffset, sSize, gSize, cSize, wPlane=None):
        self.ctx = ctx
        self.iu = iu if isinstance(iu, DeviceArray) else ctx.to_device(np.asarray(iu, np.int32))
        self.iv = iv if isinstance(iv, DeviceArray) else ctx.to_device(np.asarray(iv, np.int32))
        self.cOffset = (cOffset if isinstance(cOffset, DeviceArray)
                        else ctx.to_device(np.asarray(cOffset, np.int32)))
        self.wPlane = None
        if wPlane is not None:
            self.wPlane = (wPlane if isinstance(wPlane, DeviceArray)
                           else ctx.to_device(np.asarray(wPlane, np.int32)))
        self.sSize = np.ascontiguousarray(np.atleast_1d(sSize), dtype=np.int32)
        self.gSize = int(gSize)
        self.cSize = int(cSize)
        self.nVis = self.iu.size

    def geometry(self) -> Geometry:
        g = Geometry()
        g.nVis = self.nVis
        g.d_iu = self.iu.ptr
        g.d_iv = self.iv.ptr
        g.d_cOffset = self.cOffset.ptr
        g.d_wPlane = self.wPlane.ptr if self.wPlane is not None else None
        g.gSize = self.gSize
        g.nPlanes = int(self.sSize.size)
        g.sSize = self.sSize.ctypes.data_as(P)
        g.cSize = self.cSize
        return g


def grid_plain(ctx: Context, vis: Visibilities, data: DeviceArray, C: DeviceArray,
               grid: DeviceArray):
    g = vis.geometry()
    ctx.check(lib().tcv_grid_plain(ctx.h, ct.byref(g), data.ptr, C.ptr, grid.ptr), "grid_plain")


def degrid_plain(ctx: Context, vis: Visibilities, grid: DeviceArray, C: DeviceArray,
                 out: DeviceArray):
    g = vis.geometry()
    ctx.check(lib().tcv_degrid_plain(ctx.h, ct.byref(g), grid.ptr, C.ptr, out.ptr),
              "degrid_plain")


def grid_plain_f64(ctx: Context, vis: Visibilities, data: DeviceArray | None, C: DeviceArray,
                   grid: DeviceArray):
    """FP64 reference gridding into a complex128 grid (data None: the PSF)."""
    if grid.dtype != np.complex128:
        raise TypeError("grid_plain_f64 accumulates into a complex128 grid")
    g = vis.geometry()
    ctx.check(lib().tcv_grid_plain_f64(ctx.h, ct.byref(g), None if data is None else data.ptr, C.ptr,
                                       grid.ptr), "grid_plain_f64")


def degrid_plain_f64(ctx: Context, vis: Visibilities, grid: DeviceArray, C: DeviceArray,
                     out: DeviceArray):
    """FP64 reference degridding of a complex64 grid into complex128 outputs."""
    if out.dtype != np.complex128 or grid.dtype != np.complex64:
        raise TypeError("degrid_plain_f64 reads a complex64 grid into complex128 outputs")
    g = vis.geometry()
    ctx.check(lib().tcv_degrid_plain_f64(ctx.h, ct.byref(g), grid.ptr, C.ptr, out.ptr),
              "degrid_plain_f64")


def autotune(ctx: Context, vis: Visibilities, data: DeviceArray, C: DeviceArray,
             grid: DeviceArray, out: DeviceArray, op: int = 3, candidates=None, reps: int = 2):
    """Times candidate tunings of the tiled path on this job and returns the
    fastest as a dict (cached in the context for plans made without tuning)."""
    g = vis.geometry()
    arr, n = None, 0
    if candidates:
        arr = (Tuning * len(candidates))(*[Tuning(c.get("tile_w", 0), c.get("tile_h", 0),
                                                  c.get("chunk", 0), c.get("slice_keys", 0),
                                                  c.get("phases", 0), c.get("xcd_split", -1),
                                                  c.get("kernel", 0), c.get("degrid_kernel", 0),
                                                  c.get("grid_phases", 0), c.get("tap_bytes", 0),
                                                  c.get("item_entries", 0), c.get("grid_cols", 0))
                                           for c in candidates])
        n = len(candidates)
    best = Tuning()
    ms = ct.c_double()
    ctx.check(lib().tcv_autotune(ctx.h, ct.byref(g), data.ptr if data is not None else None, C.ptr,
                                 grid.ptr, out.ptr if out is not None else None, op, arr, n, reps,
                                 ct.byref(best), ct.byref(ms)), "autotune")
    return dict(tile_w=best.tile_w, tile_h=best.tile_h, chunk=best.chunk,
                slice_keys=best.slice_keys, phases=best.phases, xcd_split=best.xcd_split,
                kernel=best.kernel, degrid_kernel=best.degrid_kernel,
                grid_phases=best.grid_phases, tap_bytes=best.tap_bytes,
                item_entries=best.item_entries, grid_cols=best.grid_cols, ms=ms.value)


class Plan:
    """Device-side binning of a job's visibilities by grid tile (tcv_plan)."""

    def __init__(self, ctx: Context, vis: Visibilities, tile_w=0, tile_h=0, chunk=0,
                 slice_keys=0, phases=0, xcd_split=-1, kernel=0, degrid_kernel=0, grid_phases=0,
                 tap_bytes=0, item_entries=0, grid_cols=0):
        """tap_bytes=4 makes the automatic choices for FP16 storage (grid_half /
        degrid_half); 0 or 8 for FP32 tables."""
        self.ctx = ctx
        self.vis = vis
        fields = (tile_w, tile_h, chunk, slice_keys, phases, xcd_split, kernel, degrid_kernel,
                  grid_phases, tap_bytes, item_entries, grid_cols)
        explicit = fields != (0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0)
        t = Tuning(*fields)
        g = vis.geometry()
        h = P()
        # no explicit tuning: the library uses the context's autotuned choice for
        # this workload class when there is one, else its built-in defaults
        ctx.check(lib().tcv_plan_create(ctx.h, ct.byref(g), ct.byref(t) if explicit else None,
                                        ct.byref(h)), "plan_create")
        self.h = h

    def build(self, vis: Visibilities | None = None):
        vis = vis or self.vis
        g = vis.geometry()
        self.ctx.check(lib().tcv_plan_build(self.ctx.h, self.h, ct.byref(g)), "plan_build")

    def info(self):
        t = Tuning()
        e, it, bad = ct.c_longlong(), ct.c_longlong(), ct.c_longlong()
        self.ctx.check(lib().tcv_plan_info(self.ctx.h, self.h, ct.byref(t), ct.byref(e),
                                           ct.byref(it), ct.byref(bad)), "plan_info")
        return dict(tile_w=t.tile_w, tile_h=t.tile_h, chunk=t.chunk, slice_keys=t.slice_keys,
                    phases=t.phases, xcd_split=t.xcd_split, kernel=t.kernel,
                    degrid_kernel=t.degrid_kernel, grid_phases=t.grid_phases,
                    tap_bytes=t.tap_bytes, item_entries=t.item_entries, grid_cols=t.grid_cols,
                    entries=e.value,
                    items=it.value,
                    out_of_range=bad.value)

    def grid(self, data: DeviceArray, C: DeviceArray, grid: DeviceArray):
        self.ctx.check(lib().tcv_grid_tiled(self.ctx.h, self.h, data.ptr, C.ptr, grid.ptr),
                       "grid_tiled")

    def degrid(self, grid: DeviceArray, C: DeviceArray, out: DeviceArray):
        self.ctx.check(lib().tcv_degrid_tiled(self.ctx.h, self.h, grid.ptr, C.ptr, out.ptr),
                       "degrid_tiled")

    def grid_half(self, data: "HalfVis", C: "HalfTable", grid: DeviceArray):
        self.ctx.check(lib().tcv_grid_tiled_half(self.ctx.h, self.h, data.buf.ptr, data.scale,
                                                 C.buf.ptr, C.scales.ptr, C.planeTaps, grid.ptr),
                       "grid_tiled_half")

    def degrid_half(self, grid: DeviceArray, C: "HalfTable", out: DeviceArray):
        self.ctx.check(lib().tcv_degrid_tiled_half(self.ctx.h, self.h, grid.ptr, C.buf.ptr,
                                                   C.scales.ptr, C.planeTaps, out.ptr),
                       "degrid_tiled_half")

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().tcv_plan_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HalfTable:
    """C(w) in FP16 storage: half (re, im) pairs with one max-abs scale per
    w-plane.  planeTaps > 0: uniform planes (attic layout); otherwise the plane
    boundaries come from planeStart (nPlanes + 1 tap offsets)."""

    def __init__(self, ctx: Context, C: DeviceArray, nPlanes: int, planeTaps: int = 0,
                 planeStart=None):
        self.ctx = ctx
        n = C.size
        self.buf = DeviceArray(ctx, n * 2, np.float16)
        self.scales = DeviceArray(ctx, nPlanes, np.float32)
        self.planeTaps = int(planeTaps) if planeStart is None else 0
        ps = None if planeStart is None else np.ascontiguousarray(planeStart, np.int64)
        host = np.empty(nPlanes, np.float32)
        ctx.check(lib().tcv_ctable_to_half(ctx.h, C.ptr, n, nPlanes, _p(ps), int(planeTaps),
                                           self.buf.ptr, self.scales.ptr, _p(host)),
                  "ctable_to_half")
        self.host_scales = host


class HalfVis:
    """Visibilities in FP16 storage (half pairs and one max-abs scale)."""

    def __init__(self, ctx: Context, data: DeviceArray):
        self.buf = DeviceArray(ctx, data.size * 2, np.float16)
        sc = ct.c_float()
        ctx.check(lib().tcv_vis_to_half(ctx.h, data.ptr, data.size, self.buf.ptr, ct.byref(sc)),
                  "vis_to_half")
        self.scale = sc.value


class Batch:
    """Many jobs in one persistent launch over a device work queue (tcv_batch).

    jobs: sequence of (plan, data, C, grid, out) with DeviceArrays (data or
    out may be None).  All plans must share one tile shape."""

    def __init__(self, ctx: Context, jobs):
        self.ctx = ctx
        self._keep = list(jobs)
        arr = (Job * len(jobs))()
        for k, (plan, data, C, grid, out) in enumerate(jobs):
            arr[k] = Job(plan.h, data.ptr if data is not None else None, C.ptr, grid.ptr,
                         out.ptr if out is not None else None)
        h = P()
        ctx.check(lib().tcv_batch_create(ctx.h, len(jobs), arr, ct.byref(h)), "batch_create")
        self.h = h

    def rebuild(self):
        self.ctx.check(lib().tcv_batch_rebuild(self.ctx.h, self.h), "batch_rebuild")

    def grid(self):
        self.ctx.check(lib().tcv_batch_grid(self.ctx.h, self.h), "batch_grid")

    def degrid(self):
        self.ctx.check(lib().tcv_batch_degrid(self.ctx.h, self.h), "batch_degrid")

    def run_graph(self, ops: int):
        """ops: 1 rebuild + 2 grid + 4 degrid; captured once, replayed after."""
        self.ctx.check(lib().tcv_batch_run_graph(self.ctx.h, self.h, ops), "batch_run_graph")

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().tcv_batch_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def attic_initC_device(ctx: Context, freq0: float, cellSize: float, wSize: int, baseline: int,
                       overSample: int = 8):
    """attic C(w) generated on the device (FP64 taps, FP64-sum normalisation).
    Returns (support, wCellSize, DeviceArray of complex64 taps)."""
    sup, wcs = ct.c_int(), ct.c_double()
    from .setup import attic_initC
    s0, _, _ = attic_initC(freq0, cellSize, wSize, baseline, overSample, table=False)
    S = 2 * s0 + 1
    C = DeviceArray(ctx, S * S * overSample * overSample * wSize, np.complex64)
    ctx.check(lib().tcv_attic_initC_device(ctx.h, freq0, cellSize, wSize, int(baseline), overSample,
                                           ct.byref(sup), ct.byref(wcs), C.ptr),
              "attic_initC_device")
    return sup.value, wcs.value, C


class Server:
    """Continuous-batching request loop on the device (tcv_server).

    Requests reference resident inputs (a Visibilities geometry, C(w), and the
    data to grid or the grid to degrid); the server's HBM arena holds each
    request's plan and result."""

    GRID, DEGRID = 0, 1

    def __init__(self, ctx: Context, arena_bytes: int, max_inflight: int = 2, max_batch: int = 0,
                 auto_release: bool = False, tile_w=0, tile_h=0, chunk=0, slice_keys=0, phases=0,
                 max_batch_points: int = 0, item_entries: int = 0, grid_cols: int = 0,
                 exclusive_points: int = 0):
        self.ctx = ctx
        cfg = ServerConfig(int(arena_bytes), max_inflight, max_batch, tile_w, tile_h, chunk,
                           slice_keys, phases, int(auto_release), int(max_batch_points),
                           int(item_entries), int(grid_cols), int(exclusive_points))
        h = P()
        ctx.check(lib().tcv_server_create(ctx.h, ct.byref(cfg), ct.byref(h)), "server_create")
        self.h = h
        self._keep = {}

    def _check(self, st, what):
        if st:
            raise TcvError(st, f"{what}: {lib().tcv_server_last_error(self.h).decode()}")

    def submit(self, op: int, vis: Visibilities, C: DeviceArray, data: DeviceArray = None,
               grid: DeviceArray = None) -> int:
        req = Request(op, vis.geometry(), data.ptr if data is not None else None, C.ptr,
                      grid.ptr if grid is not None else None)
        rid = ct.c_longlong()
        self._check(lib().tcv_server_submit(self.h, ct.byref(req), ct.byref(rid)), "server_submit")
        self._keep[rid.value] = vis   # the geometry's device arrays must outlive the request
        return rid.value

    def footprint(self, op: int, vis: Visibilities, C: DeviceArray, data: DeviceArray = None,
                  grid: DeviceArray = None) -> int:
        """Arena bytes the request would hold while it runs (plan + result)."""
        req = Request(op, vis.geometry(), data.ptr if data is not None else None, C.ptr,
                      grid.ptr if grid is not None else None)
        b = ct.c_size_t()
        self._check(lib().tcv_server_footprint(self.h, ct.byref(req), ct.byref(b)), "server_footprint")
        return b.value

    def cancel(self, rid: int):
        self._check(lib().tcv_server_cancel(self.h, rid), "server_cancel")

    def wait(self, rid: int, timeout: float = 120.0) -> dict:
        r = Result()
        self._check(lib().tcv_server_wait(self.h, rid, timeout, ct.byref(r)), "server_wait")
        return dict(submitted=r.submitted, admitted=r.admitted, launched=r.launched,
                    completed=r.completed, status=r.status, d_result=r.d_result, points=r.points)

    def download(self, res: dict, shape, dtype=np.complex64) -> np.ndarray:
        out = np.empty(shape, dtype)
        self.ctx.check(lib().tcv_download(self.ctx.h, _p(out), res["d_result"], out.nbytes),
                       "server result download")
        return out

    def release(self, rid: int):
        self._check(lib().tcv_server_release(self.h, rid), "server_release")
        self._keep.pop(rid, None)

    def stats(self) -> dict:
        q, r = ct.c_longlong(), ct.c_longlong()
        u, c = ct.c_size_t(), ct.c_size_t()
        self._check(lib().tcv_server_stats(self.h, ct.byref(q), ct.byref(r), ct.byref(u),
                                           ct.byref(c)), "server_stats")
        return dict(queued=q.value, running=r.value, arena_used=u.value, arena_capacity=c.value)

    def clock(self) -> float:
        return lib().tcv_server_clock(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().tcv_server_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
