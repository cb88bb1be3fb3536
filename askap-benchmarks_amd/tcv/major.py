"""The imaging major cycle around the gridder, device-resident
(current/tMajorACC/tMajorACC.cc): PSF and image gridding, the shifted inverse
FFT with fftFix, Hogbom CLEAN, the forward FFT of the model, degridding and the
visibility subtraction.  Function names follow the reference's free functions;
every call runs in libtcv on the device.
"""
from __future__ import annotations

import ctypes as ct
import time

import numpy as np

from ._lib import lib
from .device import Context, DeviceArray, Plan, Visibilities
from .gridder import init_offsets
from .setup import LCG, attic_freq, attic_initC, attic_uvw

MAXINT = 2147483647


def gridKernel(ctx: Context, plan: Plan, data: DeviceArray | None, C: DeviceArray,
               grid: DeviceArray, isPSF: bool = False):
    """gridKernel(..., isPSF) (tMajorACC.cc:86-126): accumulates into grid;
    with isPSF every visibility counts as 1 + 0i."""
    if isPSF:
        ctx.check(lib().tcv_grid_psf(ctx.h, plan.h, C.ptr, grid.ptr), "grid_psf")
    else:
        plan.grid(data, C, grid)


def degridKernel(ctx: Context, plan: Plan, grid: DeviceArray, C: DeviceArray, data: DeviceArray):
    """degridKernel (tMajorACC.cc:239-338): overwrites data."""
    plan.degrid(grid, C, data)


def fftExec(ctx: Context, grid: DeviceArray, gSize: int, forward: bool):
    """fftExec / fftExecGPU (tMajorACC.cc:488-618): in place, unnormalised."""
    ctx.check(lib().tcv_fft_exec(ctx.h, grid.ptr, gSize, 1 if forward else 0), "fft_exec")


def fftFix(ctx: Context, grid: DeviceArray, scale: float):
    """fftFix (tMajorACC.cc:620-636): real part times scale, imaginary part 0."""
    ctx.check(lib().tcv_fft_fix(ctx.h, grid.ptr, grid.size, ct.c_float(scale)), "fft_fix")


def fftExecFix(ctx: Context, grid: DeviceArray, gSize: int, forward: bool, scale: float):
    """fftExec then fftFix with the fix folded into the transform's last pass."""
    ctx.check(lib().tcv_fft_exec_fix(ctx.h, grid.ptr, gSize, 1 if forward else 0,
                                     ct.c_float(scale)), "fft_exec_fix")


def findPeak(ctx: Context, image: DeviceArray):
    """findPeak (tMajorACC.cc:689-701): (signed real value, flat position)."""
    v, p = ct.c_float(), ct.c_longlong()
    ctx.check(lib().tcv_find_peak(ctx.h, image.ptr, image.size, ct.byref(v), ct.byref(p)),
              "find_peak")
    return v.value, p.value


def deconvolve(ctx: Context, residual: DeviceArray, dirtyWidth: int, psf: DeviceArray,
               psfWidth: int, model: DeviceArray, niters: int, gain: float = 0.1,
               peaks: bool = True):
    """deconvolve (Hogbom, tMajorACC.cc:800-840) in place on residual and
    model.  Returns ((psfPeakVal, psfPeakPos), (dirtyPeakVal, dirtyPeakPos))
    when peaks is true (the call then waits for the loop), else None."""
    if peaks:
        pv, dv = ct.c_float(), ct.c_float()
        pp, dp = ct.c_longlong(), ct.c_longlong()
        ctx.check(lib().tcv_hogbom(ctx.h, residual.ptr, dirtyWidth, psf.ptr, psfWidth, model.ptr,
                                   int(niters), ct.c_float(gain), ct.byref(pv), ct.byref(pp),
                                   ct.byref(dv), ct.byref(dp)), "hogbom")
        return (pv.value, pp.value), (dv.value, dp.value)
    ctx.check(lib().tcv_hogbom(ctx.h, residual.ptr, dirtyWidth, psf.ptr, psfWidth, model.ptr,
                               int(niters), ct.c_float(gain), None, None, None, None), "hogbom")
    return None


def subtractVis(ctx: Context, data: DeviceArray, model: DeviceArray):
    """data = data - model (tMajorACC.cc:1549-1556)."""
    ctx.check(lib().tcv_vis_subtract(ctx.h, data.ptr, model.ptr, data.size), "vis_subtract")


def major_initC(support: int, overSample: int, wSize: int) -> np.ndarray:
    """tMajorACC's initC table (tMajorACC.cc:388-424): only tap (i=0, j=0) of
    each (osi, osj, w) slice is set (the cos/exp assignment is commented out in
    the reference), and normalising by the float sum of |C| over the table
    leaves those taps at exactly 1."""
    s = 2 * support + 1
    C = np.zeros(s * s * overSample * overSample * wSize, np.complex64)
    C[:: s * s] = 1.0
    return C


class MajorCycle:
    """tMajorACC's main loop (tMajorACC.cc:889-1558) resident on one device.

    Set-up as the reference: the LCG uvw (baseline 2000 m by default), the
    attic frequencies, support / wCellSize from the attic formula, the
    delta-slice C table, a true sky of nSources point sources drawn from the
    same LCG, transformed by fftExec and degridded into the visibilities.
    Each cycle: (cycle 0) grid the PSF, inverse FFT + fftFix(1/nVis); grid the
    data, inverse FFT + fftFix(1/nVis) into the dirty image; Hogbom nMinor
    iterations into a zeroed model image; forward FFT of the model; degrid
    it; subtract the model visibilities from the data.
    """

    def __init__(self, ctx: Context, nSamples: int = 160000, wSize: int = 33, nChan: int = 1,
                 cellSize: float = 5.0, gSize: int = 4096, baseline: int = 2000,
                 nMajor: int = 5, nMinor: int = 100, nSources: int = 100, gain: float = 0.1,
                 overSample: int = 8, tuning: dict | None = None):
        self.ctx = ctx
        self.gSize, self.nMajor, self.nMinor, self.gain = gSize, nMajor, nMinor, gain
        lcg = LCG()
        u, v, w = attic_uvw(nSamples, baseline, lcg)
        freq = attic_freq(nChan)
        support, wCellSize, _ = attic_initC(freq[0], cellSize, wSize, baseline, overSample,
                                            table=False)
        self.support = support
        s = 2 * support + 1
        C = major_initC(support, overSample, wSize)
        self.C = ctx.to_device(C)
        du, dv, dw = ctx.to_device(u), ctx.to_device(v), ctx.to_device(w)
        iu, iv, _, co, _, bad = init_offsets(ctx, du, dv, dw, freq, cellSize, wCellSize, wSize,
                                             gSize, overSample, [s], None, want_wplane=False)
        if bad:
            raise ValueError(f"{bad} visibilities fall outside the {gSize}^2 grid")
        self.nVis = nSamples * nChan
        self.vis = Visibilities(ctx, iu, iv, co, [s], gSize, C.size)
        self.plan = Plan(ctx, self.vis, **(tuning or {}))
        # true sky (tMajorACC.cc:975-983): l, m from the LCG, flux r / maxint
        sky = np.zeros(gSize * gSize, np.complex64)
        for _ in range(nSources):
            l = int(gSize * float(lcg.randomInt()) / float(MAXINT))
            m = int(gSize * float(lcg.randomInt()) / float(MAXINT))
            sky[m * gSize + l] += np.float32(lcg.randomInt()) / np.float32(MAXINT)
        self.sky = sky.reshape(gSize, gSize)
        true_grid = ctx.to_device(self.sky)
        fftExec(ctx, true_grid, gSize, True)
        self.data = ctx.empty(self.nVis, np.complex64)
        degridKernel(ctx, self.plan, true_grid, self.C, self.data)
        true_grid.free()
        self.model_vis = ctx.zeros(self.nVis, np.complex64)
        self.psf = ctx.empty((gSize, gSize), np.complex64)
        self.img = ctx.empty((gSize, gSize), np.complex64)
        self.model = ctx.empty((gSize, gSize), np.complex64)
        self.peaks = []
        self.times = {}

    def _t(self, name, t0):
        self.ctx.sync()
        self.times[name] = self.times.get(name, 0.0) + time.perf_counter() - t0

    def cycle(self, it: int, timed: bool = False):
        ctx, g = self.ctx, self.gSize
        scale = 1.0 / float(self.nVis)
        mark = (lambda name, t0: self._t(name, t0)) if timed else (lambda name, t0: None)
        if it == 0:
            t0 = time.perf_counter()
            self.psf.zero()
            gridKernel(ctx, self.plan, None, self.C, self.psf, isPSF=True)
            mark("psf_grid", t0)
        t0 = time.perf_counter()
        self.img.zero()
        gridKernel(ctx, self.plan, self.data, self.C, self.img)
        mark("grid", t0)
        if it == 0:
            t0 = time.perf_counter()
            fftExecFix(ctx, self.psf, g, False, scale)
            mark("psf_ifft", t0)
        t0 = time.perf_counter()
        fftExecFix(ctx, self.img, g, False, scale)
        mark("ifft", t0)
        t0 = time.perf_counter()
        self.model.zero()
        pk = deconvolve(ctx, self.img, g, self.psf, g, self.model, self.nMinor, self.gain,
                        peaks=(it == 0 or timed))
        if pk:
            self.peaks.append(pk)
        mark("hogbom", t0)
        t0 = time.perf_counter()
        fftExec(ctx, self.model, g, True)
        mark("fft", t0)
        t0 = time.perf_counter()
        degridKernel(ctx, self.plan, self.model, self.C, self.model_vis)
        mark("degrid", t0)
        t0 = time.perf_counter()
        subtractVis(ctx, self.data, self.model_vis)
        mark("subtract", t0)

    def run(self, timed: bool = False):
        for it in range(self.nMajor):
            self.cycle(it, timed)
        self.ctx.sync()
