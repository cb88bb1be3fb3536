"""Drop-in `Benchmark` classes with the reference's method names, running the
hot path (initCOffset, gridKernel, degridKernel) on the GPU through libtcv.

  Benchmark         attic/tConvolveMPI/Benchmark.h:55-99  (uniform support,
                    real C, compile-time constants become constructor args)
  CurrentBenchmark  current/tConvolveMPI/Benchmark.h:38-113 (runType 0..4,
                    per-plane widths, complex C)

Arrays named like the reference members (grid, C, samples) live on the device;
`grid_host()`, `outdata_host()` and `samples_host()` download them.
"""
from __future__ import annotations

import numpy as np

from .device import Context, DeviceArray, Plan, Visibilities, degrid_plain, grid_plain
from .gridder import init_offsets
from .setup import (LCG, attic_freq, attic_initC, current_initC, current_params, current_uvw,
                    current_wavenumber)


class Benchmark:
    """attic/tConvolveMPI/Benchmark.h:55-99 on one MI355X.

    method "tiled" (default) bins visibilities by grid tile on the device;
    method "plain" uses the one-workgroup-per-visibility atomic kernels."""

    def __init__(self, ctx: Context | None = None, nSamples: int = 160000, wSize: int = 33,
                 nChan: int = 1, gSize: int = 4096, cellSize: float = 5.0, baseline: int = 2000,
                 method: str = "tiled", tuning: dict | None = None):
        self.ctx = ctx or Context(0)
        self.nSamples, self.wSize, self.nChan = nSamples, wSize, nChan
        self.gSize, self.cellSize, self.baseline = gSize, cellSize, baseline
        self.method = method
        self.tuning = tuning or {}
        self._lcg = LCG(1)          # Benchmark::Benchmark() : next(1)
        self.m_support = 0
        self.overSample = 0
        self.wCellSize = 0.0
        self.plan = None

    # attic/tConvolveMPI/Benchmark.cc:53-58
    def randomInt(self) -> int:
        return self._lcg.randomInt()

    # attic/tConvolveMPI/Benchmark.cc:60-96
    def init(self):
        n = self.nSamples
        u = np.empty(n)
        v = np.empty(n)
        w = np.empty(n)
        maxint = 2147483647.0
        # three draws per sample in u, v, w order, integer baseline / 2
        for i in range(n):
            u[i] = self.baseline * float(self.randomInt()) / maxint - self.baseline // 2
            v[i] = self.baseline * float(self.randomInt()) / maxint - self.baseline // 2
            w[i] = self.baseline * float(self.randomInt()) / maxint - self.baseline // 2
        self.u, self.v, self.w = u, v, w
        self.data = self.ctx.to_device(np.ones(n * self.nChan, np.complex64))
        self.outdata = self.ctx.zeros(n * self.nChan, np.complex64)
        self.grid = self.ctx.zeros((self.gSize, self.gSize), np.complex64)
        freq = attic_freq(self.nChan)
        self.m_support, self.overSample, self.wCellSize, C = self.initC(freq, self.cellSize,
                                                                       self.wSize)
        self.C = self.ctx.to_device(C)
        self.initCOffset(u, v, w, freq, self.cellSize, self.wCellSize, self.wSize, self.gSize,
                         self.m_support, self.overSample)

    def runGrid(self):
        self.gridKernel(self.m_support, self.C, self.grid, self.gSize)

    def runDegrid(self):
        self.degridKernel(self.grid, self.gSize, self.m_support, self.C, self.outdata)

    # attic/tConvolveMPI/Benchmark.cc:125-155 (accumulates into grid)
    def gridKernel(self, support: int, C: DeviceArray, grid: DeviceArray, gSize: int):
        self._check_support(support, gSize)
        if self.method == "tiled":
            self._plan().grid(self.data, C, grid)
        else:
            grid_plain(self.ctx, self.samples, self.data, C, grid)

    # attic/tConvolveMPI/Benchmark.cc:158-195 (overwrites data)
    def degridKernel(self, grid: DeviceArray, gSize: int, support: int, C: DeviceArray,
                     data: DeviceArray):
        self._check_support(support, gSize)
        if self.method == "tiled":
            self._plan().degrid(grid, C, data)
        else:
            degrid_plain(self.ctx, self.samples, grid, C, data)

    # attic/tConvolveMPI/Benchmark.cc:206-270; returns (support, overSample, wCellSize, C)
    def initC(self, freq, cellSize: float, wSize: int):
        support, wcs, C = attic_initC(float(freq[0]), cellSize, wSize, self.baseline, 8)
        return support, 8, wcs, C

    # attic/tConvolveMPI/Benchmark.cc:281-324, on the device in FP64
    def initCOffset(self, u, v, w, freq, cellSize, wCellSize, wSize, gSize, support, overSample):
        ctx = self.ctx
        du, dv, dw = ctx.to_device(u), ctx.to_device(v), ctx.to_device(w)
        s = 2 * support + 1
        iu, iv, _, co, _, bad = init_offsets(ctx, du, dv, dw, freq, cellSize, wCellSize, wSize,
                                             gSize, overSample, [s], None, want_wplane=False,
                                             allow_out_of_range=True)
        self.nOutOfRange = bad
        self.samples = Visibilities(ctx, iu, iv, co, [s], gSize, s * s * overSample ** 2 * wSize)
        self.plan = None

    def getSupport(self) -> int:
        return self.m_support

    # ---- helpers ---------------------------------------------------------
    def _check_support(self, support, gSize):
        if 2 * support + 1 != int(self.samples.sSize[0]) or gSize != self.samples.gSize:
            raise ValueError("support/gSize differ from the initialised problem")

    def _plan(self) -> Plan:
        if self.plan is None:
            self.plan = Plan(self.ctx, self.samples, **self.tuning)
        return self.plan

    def grid_host(self) -> np.ndarray:
        return self.grid.download()

    def outdata_host(self) -> np.ndarray:
        return self.outdata.download()

    def samples_host(self):
        return self.samples.iu.download(), self.samples.iv.download(), self.samples.cOffset.download()


def sort_by_plane(iu, iv, wPlane, cOffset):
    """setSort(1): the index arrays reordered by w-plane, keeping the input
    order within each plane (current/tConvolveMPI/Benchmark.cc:275-298, a
    counting sort; here a stable sort by plane, the same permutation)."""
    order = np.argsort(np.asarray(wPlane), kind="stable")
    return tuple(np.ascontiguousarray(np.asarray(a)[order]) for a in (iu, iv, wPlane, cOffset))


class CurrentBenchmark:
    """current/tConvolveMPI/Benchmark.h:38-113 on one MI355X."""

    def __init__(self, ctx: Context | None = None, method: str = "tiled",
                 tuning: dict | None = None):
        self.ctx = ctx or Context(0)
        self.method = method
        self.tuning = tuning or {}
        self._lcg = LCG(1)
        self.mpirank = 0
        self.doSort = 0
        self.runType = 0
        self.plan = None

    def randomInt(self) -> int:
        return self._lcg.randomInt()

    def setMPIrank(self, rank: int):
        self.mpirank = rank

    def setSort(self, t: int):
        self.doSort = t

    def setRunType(self, t: int):
        self.runType = t

    def getRunType(self) -> int:
        return self.runType

    def getSupport(self) -> int:
        return self.m_support

    # current/tConvolveMPI/Benchmark.cc:61-300
    def init(self):
        p = current_params(self.runType)
        self.gSize, self.uvCellSize, self.wSize = p["gSize"], p["uvCellSize"], p["wSize"]
        self.m_support, self.overSample, self.wCellSize = p["support"], p["overSample"], p["wCellSize"]
        self.baseline = p["baseline"]
        self.u, self.v, self.w = current_uvw(self.baseline, self._lcg)
        self.nSamples, self.nChan = self.u.size, 1
        self.sSize, self.cOffset0, C = current_initC(self.uvCellSize, self.wSize, self.m_support,
                                                     self.overSample, self.wCellSize)
        self.C = self.ctx.to_device(C)
        self.data = self.ctx.to_device(np.ones(self.nSamples, np.complex64))
        self.outdata1 = self.ctx.zeros(self.nSamples, np.complex64)
        self.grid1 = self.ctx.zeros((self.gSize, self.gSize), np.complex64)
        wn = current_wavenumber(1)
        ctx = self.ctx
        du, dv, dw = ctx.to_device(self.u), ctx.to_device(self.v), ctx.to_device(self.w)
        iu, iv, wp, co, npp, bad = init_offsets(ctx, du, dv, dw, wn, self.uvCellSize,
                                                self.wCellSize, self.wSize, self.gSize,
                                                self.overSample, self.sSize, self.cOffset0,
                                                want_wplane=True, allow_out_of_range=True)
        self.numPerPlane = npp
        self.nOutOfRange = bad
        if self.doSort == 1 and self.wSize > 1:
            # like the reference, the data vector is not permuted
            h = sort_by_plane(*(a.download() for a in (iu, iv, wp, co)))
            iu, iv, wp, co = (ctx.to_device(a) for a in h)
        self.samples = Visibilities(ctx, iu, iv, co, self.sSize, self.gSize, C.size, wPlane=wp)
        self.plan = None

    def runGrid(self):
        self.gridKernel(self.C, self.grid1, self.gSize)

    def runDegrid(self):
        self.degridKernel(self.grid1, self.gSize, self.C, self.outdata1)

    # current/tConvolveMPI/Benchmark.cc:373-413
    def gridKernel(self, C: DeviceArray, grid: DeviceArray, gSize: int):
        if self.method == "tiled":
            self._plan().grid(self.data, C, grid)
        else:
            grid_plain(self.ctx, self.samples, self.data, C, grid)

    # current/tConvolveMPI/Benchmark.cc:416-460
    def degridKernel(self, grid: DeviceArray, gSize: int, C: DeviceArray, data: DeviceArray):
        if self.method == "tiled":
            self._plan().degrid(grid, C, data)
        else:
            degrid_plain(self.ctx, self.samples, grid, C, data)

    def nVisibilitiesGridded(self) -> int:
        return self.nSamples * self.nChan

    # current/tConvolveMPI/Benchmark.cc:683-700
    def nPixelsGridded(self) -> int:
        npp = self.numPerPlane.astype(np.int64)
        if int(npp.sum()) != self.nVisibilitiesGridded():
            return -1
        s = self.sSize.astype(np.int64)
        return int((npp * s * s).sum())

    # current/tConvolveMPI/Benchmark.cc:702-740 (ASKAP requirement model, per process)
    @staticmethod
    def requiredRate():
        nvis = int((36 * 35) / 2 * 12 * 3600. / 5.)
        cont = np.float32(10 * nvis * 10 * 3 * 3 * 1 * 1 * 400 / 5250.)
        nvis = int((30 * 29) / 2 * 12 * 3600. / 5.)
        spec = np.float32(10 * nvis * 4 * 3 * 1 * 1 * 78 * 100 / 5500.)
        return [float(cont), float(spec)]

    def _plan(self) -> Plan:
        if self.plan is None:
            self.plan = Plan(self.ctx, self.samples, **self.tuning)
        return self.plan

    def grid_host(self) -> np.ndarray:
        return self.grid1.download()

    def outdata_host(self) -> np.ndarray:
        return self.outdata1.download()
