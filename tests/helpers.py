"""Shared test helpers: seeded small problems and the error metrics."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    arrs = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return meta, arrs


def rel_err(a, ref):
    """max|a - ref| / max|ref| separately on real and imaginary parts (the
    larger of the two).  A part whose reference is identically zero must match
    exactly (returns inf otherwise)."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    worst = 0.0
    for f in (np.real, np.imag):
        r = f(ref).astype(np.float64)
        d = np.abs(f(a).astype(np.float64) - r).max() if r.size else 0.0
        m = np.abs(r).max() if r.size else 0.0
        if m == 0.0:
            if d != 0.0:
                return float("inf")
            continue
        worst = max(worst, d / m)
    return worst


def random_C(rng, S, os_, nPlanes, complex_=True):
    """A C table in the attic layout [plane][osj][osi][S][S] with taps of
    unit-ish magnitude and both signs."""
    n = S * S * os_ * os_ * nPlanes
    re = rng.standard_normal(n).astype(np.float32)
    im = rng.standard_normal(n).astype(np.float32) if complex_ else np.zeros(n, np.float32)
    return (re + 1j * im).astype(np.complex64) * np.float32(1.0 / S)


def small_problem(seed, nVis, gSize, S, os_=8, nPlanes=5, complex_=True, margin=0,
                  cluster=None):
    """Random geometry-level inputs: iu, iv, cOffset (attic layout), data, C."""
    rng = np.random.default_rng(seed)
    half = S // 2
    # valid centres: columns iu in [half, gSize-1-half] (footprint centred on
    # iu), rows iv in [0, gSize-S] (footprint starts at row iv, as in the
    # reference's gind = iu + gSize*iv - support)
    ulo, uhi = half + margin, gSize - half - 1 - margin
    vlo, vhi = margin, gSize - S - margin
    if cluster:
        cu = rng.integers(ulo + cluster, uhi - cluster)
        cv = rng.integers(vlo + cluster, vhi - cluster)
        iu = np.clip(cu + rng.integers(-cluster, cluster + 1, nVis), ulo, uhi)
        iv = np.clip(cv + rng.integers(-cluster, cluster + 1, nVis), vlo, vhi)
    else:
        iu = rng.integers(ulo, uhi + 1, nVis)
        iv = rng.integers(vlo, vhi + 1, nVis)
    slices = rng.integers(0, os_ * os_ * nPlanes, nVis)
    cOffset = (slices * S * S).astype(np.int32)
    data = ((rng.standard_normal(nVis) + 1j * rng.standard_normal(nVis)) *
            np.sqrt(0.5)).astype(np.complex64)
    C = random_C(rng, S, os_, nPlanes, complex_)
    return iu.astype(np.int32), iv.astype(np.int32), cOffset, data, C


def exact_problem(seed, nVis, gSize, Sz, cluster):
    """Inputs whose every partial sum is exact in FP32: C taps are multiples
    of 1/4 in [-1, 1], data and grid values small integers.  Results then do
    not depend on summation order, so different kernels, plans and the oracle
    must agree bit for bit."""
    iu, iv, co, _, C = small_problem(seed, nVis, gSize, Sz, cluster=cluster)
    rng = np.random.default_rng(seed + 1)
    C = (rng.integers(-4, 5, C.size) / 4 + 1j * rng.integers(-4, 5, C.size) / 4).astype(np.complex64)
    data = (rng.integers(-3, 4, nVis) + 1j * rng.integers(-3, 4, nVis)).astype(np.complex64)
    grid = (rng.integers(-2, 3, (gSize, gSize)) + 1j * rng.integers(-2, 3, (gSize, gSize))).astype(np.complex64)
    return iu, iv, co, data, C, grid


def grid_cols_choices(tw, th):
    """Column-group widths the pair gridder is built with for a tile shape:
    the tile width, and half and a quarter of it (at least 16) on tiles 64
    wide and up whose height gives every wave whole row groups."""
    out = [tw]
    if tw >= 64:
        out += [c for c in (tw // 2, tw // 4) if c >= 16 and th >= 512 // c]
    return out
