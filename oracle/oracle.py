"""ctypes binding for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the measured or
shipped path.  Complex arrays are numpy complex64 (interleaved float pairs,
the layout of std::complex<float>).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> None:
    """Compile liboracle.so (gcc, the reference's flags plus -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ct.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ct.c_void_p
I = ct.c_int
D = ct.c_double


def _declare(L):
    L.orc_random_int.restype = I
    L.orc_random_int.argtypes = [ct.POINTER(ct.c_uint64)]
    L.orc_attic_uvw.argtypes = [I, I, ct.POINTER(ct.c_uint64), P, P, P]
    L.orc_attic_freq.argtypes = [I, P]
    L.orc_attic_initC.argtypes = [D, D, I, I, I, ct.POINTER(I), ct.POINTER(D), P]
    L.orc_attic_initCOffset.argtypes = [I, P, P, P, I, P, D, D, I, I, I, I, P, P, P]
    L.orc_current_initC.restype = ct.c_long
    L.orc_current_initC.argtypes = [D, I, I, I, D, P, P, P]
    L.orc_current_initCOffset.argtypes = [I, P, P, P, I, P, D, D, I, I, I, P, P, P, P, P, P, P]
    L.orc_current_params.restype = I
    L.orc_current_params.argtypes = [I, ct.POINTER(I), ct.POINTER(D), ct.POINTER(I),
                                      ct.POINTER(I), ct.POINTER(I), ct.POINTER(D),
                                      ct.POINTER(ct.c_float)]
    L.orc_current_uvw.restype = I
    L.orc_current_uvw.argtypes = [ct.c_float, ct.POINTER(ct.c_uint64), P, P, P]
    L.orc_grid.argtypes = [I, P, P, P, P, P, P, I, P, P, I]
    L.orc_degrid.argtypes = [I, P, I, P, P, P, P, P, I, P, P]
    L.orc_grid_f64.argtypes = [I, P, P, P, P, P, P, I, P, P, I]
    L.orc_degrid_f64.argtypes = [I, P, I, P, P, P, P, P, I, P, P]
    L.orc_grid_points.restype = ct.c_longlong
    L.orc_grid_points.argtypes = [I, P, P, I]
    L.orc_sort_by_plane.argtypes = [I, I, P, P, P, P, P]
    L.orc_grid_window.argtypes = [I, P, P, P, P, P, P, I, P, I, I, I, I, P, I]
    L.orc_cpu_baseline.restype = I
    L.orc_cpu_baseline.argtypes = [I, I, I, P, P, P, P, P, P, I, P, I,
                                   ct.POINTER(D), ct.POINTER(D)]
    LL = ct.c_longlong
    L.orc_fft_fix.argtypes = [P, LL, ct.c_float]
    L.orc_find_peak.argtypes = [P, LL, ct.POINTER(ct.c_float), ct.POINTER(LL)]
    L.orc_subtract_psf.argtypes = [P, LL, P, LL, LL, LL, ct.c_float, ct.c_float]
    L.orc_deconvolve.argtypes = [P, LL, P, LL, P, I, ct.c_float, ct.POINTER(ct.c_float),
                                 ct.POINTER(LL), ct.POINTER(ct.c_float), ct.POINTER(LL)]
    L.orc_vis_subtract.argtypes = [P, P, LL]


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def _c(a, dtype):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dtype)


# ---- set-up -----------------------------------------------------------------

class LCG:
    """Benchmark::randomInt state (seed next = 1)."""

    def __init__(self, seed: int = 1):
        self.state = ct.c_uint64(seed)

    def next(self) -> int:
        return lib().orc_random_int(ct.byref(self.state))


def attic_uvw(n, baseline=2000, lcg: LCG | None = None):
    lcg = lcg or LCG()
    u = np.empty(n, np.float64)
    v = np.empty(n, np.float64)
    w = np.empty(n, np.float64)
    lib().orc_attic_uvw(n, baseline, ct.byref(lcg.state), _p(u), _p(v), _p(w))
    return u, v, w


def attic_freq(nChan):
    f = np.empty(nChan, np.float64)
    lib().orc_attic_freq(nChan, _p(f))
    return f


def attic_initC(freq0, cellSize, wSize, baseline, overSample=8, table=True):
    sup = I()
    wcs = D()
    lib().orc_attic_initC(freq0, cellSize, wSize, baseline, overSample,
                          ct.byref(sup), ct.byref(wcs), None)
    C = None
    if table:
        s = 2 * sup.value + 1
        C = np.empty(s * s * overSample * overSample * wSize, np.complex64)
        lib().orc_attic_initC(freq0, cellSize, wSize, baseline, overSample,
                              ct.byref(sup), ct.byref(wcs), _p(C))
    return sup.value, wcs.value, C


def attic_initCOffset(u, v, w, freq, cellSize, wCellSize, wSize, gSize, support,
                      overSample=8):
    u, v, w, freq = (_c(x, np.float64) for x in (u, v, w, freq))
    n = u.size * freq.size
    iu = np.empty(n, np.int32)
    iv = np.empty(n, np.int32)
    co = np.empty(n, np.int32)
    lib().orc_attic_initCOffset(u.size, _p(u), _p(v), _p(w), freq.size, _p(freq),
                                cellSize, wCellSize, wSize, gSize, support,
                                overSample, _p(iu), _p(iv), _p(co))
    return iu, iv, co


def current_params(runType):
    g, ws, sup, os_ = I(), I(), I(), I()
    uvc, wcs = D(), D()
    bl = ct.c_float()
    rc = lib().orc_current_params(runType, ct.byref(g), ct.byref(uvc), ct.byref(ws),
                                  ct.byref(sup), ct.byref(os_), ct.byref(wcs),
                                  ct.byref(bl))
    if rc:
        raise ValueError(f"unsupported runType {runType}")
    return dict(gSize=g.value, uvCellSize=uvc.value, wSize=ws.value,
                support=sup.value, overSample=os_.value, wCellSize=wcs.value,
                baseline=bl.value)


def current_uvw(baseline, lcg: LCG | None = None):
    lcg = lcg or LCG()
    n = lib().orc_current_uvw(baseline, ct.byref(lcg.state), None, None, None)
    u = np.empty(n, np.float64)
    v = np.empty(n, np.float64)
    w = np.empty(n, np.float64)
    lib().orc_current_uvw(baseline, ct.byref(lcg.state), _p(u), _p(v), _p(w))
    return u, v, w


def current_initC(uvCellSize, wSize, support, overSample, wCellSize, table=True):
    sSize = np.zeros(wSize, np.int32)
    cOffset0 = np.zeros(wSize, np.int32)
    n = lib().orc_current_initC(uvCellSize, wSize, support, overSample, wCellSize,
                                _p(sSize), _p(cOffset0), None)
    C = None
    if table:
        C = np.empty(n, np.complex64)
        lib().orc_current_initC(uvCellSize, wSize, support, overSample, wCellSize,
                                _p(sSize), _p(cOffset0), _p(C))
    return sSize, cOffset0, C


def current_initCOffset(u, v, w, wavenumber, uvCellSize, wCellSize, wSize, gSize,
                        overSample, sSize, cOffset0):
    u, v, w, wavenumber = (_c(x, np.float64) for x in (u, v, w, wavenumber))
    sSize = _c(sSize, np.int32)
    cOffset0 = _c(cOffset0, np.int32)
    n = u.size * wavenumber.size
    iu = np.empty(n, np.int32)
    iv = np.empty(n, np.int32)
    wp = np.empty(n, np.int32)
    co = np.empty(n, np.int32)
    npp = np.zeros(wSize, np.int64)
    lib().orc_current_initCOffset(u.size, _p(u), _p(v), _p(w), wavenumber.size,
                                  _p(wavenumber), uvCellSize, wCellSize, wSize, gSize,
                                  overSample, _p(sSize), _p(cOffset0), _p(iu), _p(iv),
                                  _p(wp), _p(co), _p(npp))
    return iu, iv, wp, co, npp


def sort_by_plane(iu, iv, wPlane, cOffset, numPerPlane):
    """setSort(1)'s counting sort by w-plane (current/tConvolveMPI/Benchmark.cc:275-298);
    returns sorted copies of iu, iv, wPlane, cOffset."""
    iu, iv, wp, co = (np.array(a, np.int32, copy=True) for a in (iu, iv, wPlane, cOffset))
    npp = _c(numPerPlane, np.int64)
    if wp.size and (wp.min() < 0 or wp.max() >= npp.size):
        raise ValueError("w-plane outside numPerPlane")
    if int(npp.sum()) != wp.size or not np.array_equal(np.bincount(wp, minlength=npp.size), npp):
        raise ValueError("numPerPlane does not count the w-planes")
    if lib().orc_sort_by_plane(wp.size, npp.size, _p(npp), _p(iu), _p(iv), _p(wp), _p(co)):
        raise MemoryError("orc_sort_by_plane")
    return iu, iv, wp, co


# ---- kernels ----------------------------------------------------------------

def _vis_args(iu, iv, cOffset, wPlane, sSize):
    return (_c(iu, np.int32), _c(iv, np.int32), _c(cOffset, np.int32),
            _c(wPlane, np.int32), _c(sSize, np.int32))


def _check_footprints(iu, iv, cOffset, wPlane, sSz, sUni, gSize, nTaps):
    """The C loops, like the reference, do no bounds checks: refuse inputs whose
    footprint (rows [iv, iv+s), columns [iu-s/2, iu-s/2+s)) leaves the grid or
    whose slice leaves C, instead of writing past the arrays."""
    if wPlane is not None:
        if wPlane.size and (wPlane.min() < 0 or wPlane.max() >= sSz.size):
            raise ValueError("w-plane index outside the sSize table")
        s = sSz[wPlane].astype(np.int64)
    else:
        s = np.full(iu.size, int(sUni), np.int64)
    u0 = iu.astype(np.int64) - s // 2
    v0 = iv.astype(np.int64)
    c0 = cOffset.astype(np.int64)
    bad = (u0 < 0) | (v0 < 0) | (u0 + s > gSize) | (v0 + s > gSize) | (c0 < 0) | (c0 + s * s > nTaps)
    if bad.any():
        raise ValueError(f"{int(bad.sum())} visibilities have footprints outside the grid or C")


def grid(data, iu, iv, cOffset, C, gSize, sSize, wPlane=None, out=None):
    """gridKernel: accumulates into `out` (complex64 gSize x gSize, created zero)."""
    data = _c(data, np.complex64)
    C = _c(C, np.complex64)
    iu, iv, cOffset, wPlane, sSz = _vis_args(iu, iv, cOffset, wPlane,
                                             sSize if wPlane is not None else None)
    sUni = 0 if wPlane is not None else int(sSize)
    if out is None:
        out = np.zeros((gSize, gSize), np.complex64)
    if out.shape != (gSize, gSize) or out.dtype != np.complex64 or not out.flags.c_contiguous:
        raise ValueError("out must be a C-contiguous complex64 gSize x gSize array")
    _check_footprints(iu, iv, cOffset, wPlane, sSz, sUni, gSize, C.size)
    lib().orc_grid(data.size, _p(data), _p(iu), _p(iv), _p(cOffset), _p(wPlane),
                   _p(sSz), sUni, _p(C), _p(out), gSize)
    return out


def grid_window(data, iu, iv, cOffset, C, gSize, sSize, x0, y0, w, h, wPlane=None, f64=False):
    """gridKernel's result on the window [x0, x0+w) x [y0, y0+h) only (an h x w
    complex array, row y - y0): the values grid() gives those cells, bit for
    bit (grid_f64()'s with f64=True), visiting only the footprint cells inside
    the window."""
    data = _c(data, np.complex64)
    C = _c(C, np.complex64)
    iu, iv, cOffset, wPlane, sSz = _vis_args(iu, iv, cOffset, wPlane,
                                             sSize if wPlane is not None else None)
    sUni = 0 if wPlane is not None else int(sSize)
    if not (0 <= x0 and 0 <= y0 and w > 0 and h > 0 and x0 + w <= gSize and y0 + h <= gSize):
        raise ValueError("window outside the grid")
    _check_footprints(iu, iv, cOffset, wPlane, sSz, sUni, gSize, C.size)
    out = np.zeros((h, w), np.complex128 if f64 else np.complex64)
    lib().orc_grid_window(data.size, _p(data), _p(iu), _p(iv), _p(cOffset), _p(wPlane),
                          _p(sSz), sUni, _p(C), int(x0), int(y0), int(w), int(h), _p(out),
                          1 if f64 else 0)
    return out


def degrid(grid_, iu, iv, cOffset, C, sSize, wPlane=None):
    grid_ = _c(grid_, np.complex64)
    C = _c(C, np.complex64)
    gSize = grid_.shape[0]
    iu, iv, cOffset, wPlane, sSz = _vis_args(iu, iv, cOffset, wPlane,
                                             sSize if wPlane is not None else None)
    sUni = 0 if wPlane is not None else int(sSize)
    out = np.empty(iu.size, np.complex64)
    _check_footprints(iu, iv, cOffset, wPlane, sSz, sUni, gSize, C.size)
    lib().orc_degrid(iu.size, _p(grid_), gSize, _p(iu), _p(iv), _p(cOffset),
                     _p(wPlane), _p(sSz), sUni, _p(C), _p(out))
    return out


def grid_f64(data, iu, iv, cOffset, C, gSize, sSize, wPlane=None):
    data = _c(data, np.complex128)
    C = _c(C, np.complex64)
    iu, iv, cOffset, wPlane, sSz = _vis_args(iu, iv, cOffset, wPlane,
                                             sSize if wPlane is not None else None)
    sUni = 0 if wPlane is not None else int(sSize)
    out = np.zeros((gSize, gSize), np.complex128)
    _check_footprints(iu, iv, cOffset, wPlane, sSz, sUni, gSize, C.size)
    lib().orc_grid_f64(data.size, _p(data), _p(iu), _p(iv), _p(cOffset), _p(wPlane),
                       _p(sSz), sUni, _p(C), _p(out), gSize)
    return out


def degrid_f64(grid_, iu, iv, cOffset, C, sSize, wPlane=None):
    grid_ = _c(grid_, np.complex128)
    C = _c(C, np.complex64)
    gSize = grid_.shape[0]
    iu, iv, cOffset, wPlane, sSz = _vis_args(iu, iv, cOffset, wPlane,
                                             sSize if wPlane is not None else None)
    sUni = 0 if wPlane is not None else int(sSize)
    out = np.empty(iu.size, np.complex128)
    _check_footprints(iu, iv, cOffset, wPlane, sSz, sUni, gSize, C.size)
    lib().orc_degrid_f64(iu.size, _p(grid_), gSize, _p(iu), _p(iv), _p(cOffset),
                         _p(wPlane), _p(sSz), sUni, _p(C), _p(out))
    return out


def grid_points(n, sSize, wPlane=None):
    if wPlane is None:
        return int(n) * int(sSize) * int(sSize)
    wPlane = _c(wPlane, np.int32)
    sSize = _c(sSize, np.int32)
    return int(lib().orc_grid_points(wPlane.size, _p(wPlane), _p(sSize), 0))


def cpu_baseline(nThreads, nVis, data, iu, iv, cOffset, C, gSize, sSize, wPlane=None):
    """tConvolveMPI-style replicated CPU run; returns (t_grid_s, t_degrid_s)."""
    data = _c(data, np.complex64)
    C = _c(C, np.complex64)
    iu, iv, cOffset, wPlane, sSz = _vis_args(iu, iv, cOffset, wPlane,
                                             sSize if wPlane is not None else None)
    sUni = 0 if wPlane is not None else int(sSize)
    _check_footprints(iu, iv, cOffset, wPlane, sSz, sUni, gSize, C.size)
    tg, td = D(), D()
    rc = lib().orc_cpu_baseline(nThreads, nVis, iu.size, _p(data), _p(iu), _p(iv),
                                _p(cOffset), _p(wPlane), _p(sSz), sUni, _p(C), gSize,
                                ct.byref(tg), ct.byref(td))
    if rc:
        raise MemoryError(f"cpu baseline failed (rc={rc})")
    return tg.value, td.value


# ---- major cycle (current/tMajorACC/tMajorACC.cc) -------------------------------

def fft_shift_copy(grid):
    """The reference's rotate-by-copy around its FFTs: out[(c + g/2) % g][r] =
    in[c][(r + g/2) % g] for every row c (tMajorACC.cc:509-520), written as
    explicit index loops over rows; even g only."""
    g = grid.shape[0]
    if g % 2:
        raise ValueError("fftExec does not support odd sizes")
    out = np.empty_like(grid)
    h = g // 2
    for col in range(g):
        colout = (col + h) % g
        out[colout, :h] = grid[col, h:]
        out[colout, h:] = grid[col, :h]
    return out


def fft_exec(grid, forward):
    """fftExec: shift, unnormalised 2-D C2C transform (FFTW sign convention:
    forward e^-i, backward e^+i), shift back -- tMajorACC.cc:488-546.  The
    transform runs in complex128 (numpy's pocketfft); returns complex128."""
    g = np.asarray(grid).astype(np.complex128)
    b = fft_shift_copy(g)
    b = np.fft.fft2(b) if forward else np.fft.ifft2(b) * (b.shape[0] * b.shape[1])
    return fft_shift_copy(b)


def fft_fix(grid, scale):
    """fftFix (tMajorACC.cc:620-627) on a complex64 array, in a copy."""
    g = _c(grid, np.complex64).copy()
    lib().orc_fft_fix(_p(g), g.size, ct.c_float(scale))
    return g


def find_peak(image):
    """findPeak (tMajorACC.cc:689-701): (signed real value, flat position)."""
    im = _c(image, np.complex64)
    v, pos = ct.c_float(), ct.c_longlong()
    lib().orc_find_peak(_p(im), im.size, ct.byref(v), ct.byref(pos))
    return v.value, pos.value


def subtract_psf(psf, residual, peakPos, psfPeakPos, absPeakVal, gain=0.1):
    """subtractPsf (tMajorACC.cc:723-750) on a copy of residual."""
    psf = _c(psf, np.complex64)
    r = _c(residual, np.complex64).copy()
    lib().orc_subtract_psf(_p(psf), psf.shape[0], _p(r), r.shape[0], int(peakPos),
                           int(psfPeakPos), ct.c_float(absPeakVal), ct.c_float(gain))
    return r


def deconvolve(residual, psf, model=None, niters=100, gain=0.1):
    """deconvolve (Hogbom, tMajorACC.cc:800-840) on copies; returns (residual,
    model, (psfPeakVal, psfPeakPos), (dirtyPeakVal, dirtyPeakPos))."""
    psf = _c(psf, np.complex64)
    r = _c(residual, np.complex64).copy()
    m = np.zeros_like(r) if model is None else _c(model, np.complex64).copy()
    pv, dv = ct.c_float(), ct.c_float()
    pp, dp = ct.c_longlong(), ct.c_longlong()
    lib().orc_deconvolve(_p(r), r.shape[0], _p(psf), psf.shape[0], _p(m), int(niters),
                         ct.c_float(gain), ct.byref(pv), ct.byref(pp), ct.byref(dv),
                         ct.byref(dp))
    return r, m, (pv.value, pp.value), (dv.value, dp.value)


def vis_subtract(data, model):
    """data - model (tMajorACC.cc:1549-1551), in a copy."""
    d = _c(data, np.complex64).copy()
    m = _c(model, np.complex64)
    lib().orc_vis_subtract(_p(d), _p(m), d.size)
    return d
