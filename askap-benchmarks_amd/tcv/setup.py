"""Host set-up of the gridding problem: the reference's data generators and
C(w) tables (computed by libtcv's host code), plus the synthetic workloads.

Workloads
  attic_problem     attic tConvolve (attic/tConvolveMPI/Benchmark.h:39-46):
                    160 000 samples x 1 channel from the 64-bit LCG, data 1+0i,
                    4096^2 grid, wSize 33, overSample 8, cellSize 5 wavelengths,
                    baseline 2000 m (u, v, w in [-1000, 1000] m as the reference
                    generates them); the declared variant 'wide' uses baseline
                    4000 m, i.e. [-2000, 2000] m.
  spectral_problem  spectral scale: 1 048 576 samples x 8 channels, u, v, w
                    uniform in [-6000, 6000] m (numpy PCG64, seeded), data complex
                    Gaussian with E|d|^2 = 1, freq (1.4e9 + c*1e6)/2.998e8, cellSize
                    15 wavelengths (support 64 by the reference formula), 4096^2
                    grid, wSize 129, overSample 8, attic C(w) with baseline 6000 m.
  uneven_batch      many independent jobs of uneven size for batching and serving.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field

import numpy as np

from ._lib import lib

P = ct.c_void_p


def _p(a):
    return a.ctypes.data_as(P)


class LCG:
    """Benchmark::randomInt (attic/tConvolveMPI/Benchmark.cc:53-58), seed 1."""

    def __init__(self, seed: int = 1):
        self.state = ct.c_uint64(seed)

    def randomInt(self) -> int:
        return lib().tcv_random_int(ct.byref(self.state))


def attic_uvw(n: int, baseline: int = 2000, lcg: LCG | None = None):
    lcg = lcg or LCG()
    u = np.empty(n, np.float64)
    v = np.empty(n, np.float64)
    w = np.empty(n, np.float64)
    lib().tcv_attic_uvw(n, int(baseline), ct.byref(lcg.state), _p(u), _p(v), _p(w))
    return u, v, w


def attic_freq(nChan: int) -> np.ndarray:
    """attic/tConvolveMPI/Benchmark.cc:86-90 (inverse wavelengths)."""
    i = np.arange(nChan, dtype=np.float64)
    return (1.4e9 - 2.0e5 * i / float(nChan)) / 2.998e8


def attic_initC(freq0: float, cellSize: float, wSize: int, baseline: int, overSample: int = 8,
                table: bool = True):
    """Benchmark::initC (attic): returns (support, wCellSize, C complex64 or None)."""
    sup = ct.c_int()
    wcs = ct.c_double()
    st = lib().tcv_attic_initC(freq0, cellSize, wSize, int(baseline), overSample,
                               ct.byref(sup), ct.byref(wcs), None)
    if st:
        raise ValueError("tcv_attic_initC: invalid arguments")
    C = None
    if table:
        s = 2 * sup.value + 1
        C = np.empty(s * s * overSample * overSample * wSize, np.complex64)
        lib().tcv_attic_initC(freq0, cellSize, wSize, int(baseline), overSample,
                              ct.byref(sup), ct.byref(wcs), _p(C))
    return sup.value, wcs.value, C


def current_params(runType: int) -> dict:
    g, ws, sup, os_ = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
    uvc, wcs = ct.c_double(), ct.c_double()
    bl = ct.c_float()
    if lib().tcv_current_params(runType, ct.byref(g), ct.byref(uvc), ct.byref(ws), ct.byref(sup),
                                ct.byref(os_), ct.byref(wcs), ct.byref(bl)):
        raise ValueError(f"unsupported runType {runType}")
    return dict(gSize=g.value, uvCellSize=uvc.value, wSize=ws.value, support=sup.value,
                overSample=os_.value, wCellSize=wcs.value, baseline=bl.value)


def current_uvw(baseline: float, lcg: LCG | None = None):
    lcg = lcg or LCG()
    n = lib().tcv_current_uvw(baseline, None, None, None, None)
    u = np.empty(n, np.float64)
    v = np.empty(n, np.float64)
    w = np.empty(n, np.float64)
    lib().tcv_current_uvw(baseline, ct.byref(lcg.state), _p(u), _p(v), _p(w))
    return u, v, w


def current_initC(uvCellSize, wSize, support, overSample, wCellSize, table=True):
    sSize = np.zeros(wSize, np.int32)
    cOffset0 = np.zeros(wSize, np.int32)
    n = ct.c_longlong()
    st = lib().tcv_current_initC(uvCellSize, wSize, support, overSample, wCellSize, _p(sSize),
                                 _p(cOffset0), None, ct.byref(n))
    if st:
        raise ValueError("tcv_current_initC: invalid arguments")
    C = None
    if table:
        C = np.empty(n.value, np.complex64)
        lib().tcv_current_initC(uvCellSize, wSize, support, overSample, wCellSize, _p(sSize),
                                _p(cOffset0), _p(C), ct.byref(n))
    return sSize, cOffset0, C


def current_wavenumber(nChan: int) -> np.ndarray:
    """current/tConvolveMPI/Benchmark.cc:266-269 (maxFreqHz = 1 GHz)."""
    i = np.arange(nChan, dtype=np.float64)
    return (float(np.float32(1.0e9)) - 2.0e5 * i / float(nChan)) / 2.998e8


@dataclass
class Problem:
    """Host description of one gridding job (inputs of init / initCOffset)."""
    name: str
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    freq: np.ndarray
    data: np.ndarray            # complex64, nSamples * nChan
    C: np.ndarray               # complex64 taps
    gSize: int
    cellSize: float
    wCellSize: float
    wSize: int
    overSample: int
    support: int
    sSize: np.ndarray           # per plane (1 entry when uniform)
    cOffset0: np.ndarray | None = None   # None: attic layout
    meta: dict = field(default_factory=dict)

    @property
    def nVis(self) -> int:
        return self.u.size * self.freq.size


def attic_problem(variant: str = "reference", nSamples: int = 160000) -> Problem:
    baseline = 2000 if variant == "reference" else 4000
    u, v, w = attic_uvw(nSamples, baseline)
    freq = attic_freq(1)
    # the C table and cell sizes always follow the 2000 m reference constants
    support, wcs, C = attic_initC(freq[0], 5.0, 33, 2000, 8)
    return Problem(f"attic-{variant}", u, v, w, freq, np.ones(nSamples, np.complex64), C,
                   4096, 5.0, wcs, 33, 8, support, np.array([2 * support + 1], np.int32),
                   meta=dict(baseline_m=baseline, seed="LCG next=1"))


def spectral_problem(nSamples: int = 1 << 20, nChan: int = 8, seed: int = 2024, gSize: int = 4096,
            wSize: int = 129, cellSize: float = 15.0, baseline: int = 6000) -> Problem:
    rng = np.random.Generator(np.random.PCG64(seed))
    half = baseline
    u = rng.uniform(-half, half, nSamples)
    v = rng.uniform(-half, half, nSamples)
    w = rng.uniform(-half, half, nSamples)
    freq = (1.4e9 + np.arange(nChan, dtype=np.float64) * 1.0e6) / 2.998e8
    n = nSamples * nChan
    data = ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.sqrt(0.5)).astype(np.complex64)
    support, wcs, C = attic_initC(freq[0], cellSize, wSize, baseline, 8)
    return Problem("spectral", u, v, w, freq, data, C, gSize, cellSize, wcs, wSize, 8, support,
                   np.array([2 * support + 1], np.int32),
                   meta=dict(seed=seed, baseline_m=baseline, rng="numpy PCG64"))


def current_problem(runType: int) -> Problem:
    """current/tConvolveMPI runType 0..4 (ASKAP layout, per-plane widths)."""
    p = current_params(runType)
    u, v, w = current_uvw(p["baseline"])
    sSize, cOffset0, C = current_initC(p["uvCellSize"], p["wSize"], p["support"],
                                       p["overSample"], p["wCellSize"])
    wn = current_wavenumber(1)
    return Problem(f"current-rt{runType}", u, v, w, wn, np.ones(u.size, np.complex64), C,
                   p["gSize"], p["uvCellSize"], p["wCellSize"], p["wSize"], p["overSample"],
                   p["support"], sSize, cOffset0, meta=dict(runType=runType))


F0 = 1.4e9 / 2.998e8


def attic_support(baseline: float, cellSize: float, freq0: float = F0) -> int:
    """attic/tConvolveMPI/Benchmark.cc:212-213."""
    return int(1.5 * np.sqrt(abs(baseline) * cellSize * freq0) / cellSize)


def tightest_cell(baseline: float, gSize: int, freq0: float = F0) -> float:
    """Smallest cellSize (wavelengths) for which every footprint of a
    visibility with |u|, |v| <= baseline stays on the grid: the pixel radius R
    plus the full kernel width (rows start at iv) must fit in gSize/2."""
    # R + 2*support + 2 <= gSize/2 with R = f*B/c, support ~ 1.5*sqrt(f*B/c)
    fb = freq0 * baseline
    y = (-3.0 + np.sqrt(9.0 + 4.0 * (gSize / 2 - 2))) / 2.0   # y = sqrt(f*B/c)
    c = fb / (y * y)
    for _ in range(10000):
        sup = attic_support(baseline, c, freq0)
        R = int(np.ceil(fb / c)) + 1
        if R + 2 * sup + 2 <= gSize // 2:
            return float(c)
        c *= 1.002
    raise ValueError("no cell size fits")


def uneven_batch(nJobs: int = 64, seed: int = 3, vis_min: float = 1e4, vis_max: float = 2e6,
            grids=(1024, 2048, 4096), wsizes=(33, 65, 129)):
    """The uneven batch: per job log-uniform visibility count, grid size and
    wSize drawn from the sets, baseline uniform integer in [1000, 6000] m, u, v, w
    uniform in [-B, B], complex Gaussian data, one channel at 1.4 GHz, cellSize
    at the tightest fit (support from the reference formula).  C(w) tables are
    generated on the device by the consumer (tcv.attic_initC_device)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    for j in range(nJobs):
        n = int(np.exp(rng.uniform(np.log(vis_min), np.log(vis_max))))
        g = int(rng.choice(grids))
        ws = int(rng.choice(wsizes))
        B = float(rng.integers(1000, 6001))   # integer metres, as initC takes them
        c = tightest_cell(B, g)
        u = rng.uniform(-B, B, n)
        v = rng.uniform(-B, B, n)
        w = rng.uniform(-B, B, n)
        data = ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.sqrt(0.5)).astype(np.complex64)
        sup = attic_support(B, c)
        jobs.append(dict(job=j, nVis=n, gSize=g, wSize=ws, baseline=B, cellSize=c, support=sup,
                         sSize=2 * sup + 1, u=u, v=v, w=w, data=data, freq=np.array([F0])))
    return jobs
