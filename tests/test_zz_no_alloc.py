"""tcv_grid_noalloc: the variant of tcv_grid_plain that keeps its binning plan
in the library's fixed device scratch array instead of allocating.  Called
through the C ABI and checked against tcv_grid_plain on the same inputs: bit
for bit on exact-arithmetic inputs (every partial sum exact, so summation
order cannot matter), within 1e-5 of the maximum on random complex inputs; two
calls in a row, the second smaller (it must not see the first call's plan
left in the array), accumulating into the grid as gridKernel does."""
from __future__ import annotations

import ctypes as ct

import numpy as np
import pytest

import tcv
from helpers import exact_problem, rel_err, small_problem

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _noalloc(ctx, vis, d_data, d_C, grid):
    g = vis.geometry()
    return tcv.lib().tcv_grid_noalloc(ctx.h, ct.byref(g), d_data.ptr, d_C.ptr, grid.ptr)


def _plain(ctx, vis, d_data, d_C, grid):
    g = vis.geometry()
    ctx.check(tcv.lib().tcv_grid_plain(ctx.h, ct.byref(g), d_data.ptr, d_C.ptr, grid.ptr), "grid_plain")


def _both(ctx, vis, data, C, gSize, grid0=None):
    d_data, d_C = ctx.to_device(data), ctx.to_device(C)
    out = []
    for fn in (_noalloc, _plain):
        g = ctx.to_device(grid0) if grid0 is not None else ctx.zeros((gSize, gSize), np.complex64)
        st = fn(ctx, vis, d_data, d_C, g)
        if st:
            ctx.check(st, fn.__name__)
        ctx.sync()
        out.append(g.download())
    return out


def per_plane_problem(seed, nVis, gSize, widths=(129, 65, 33), os_=8):
    """Per-plane kernel widths (the current generation's layout): planes
    concatenated at cOffset0[w], os^2 slices of s_w^2 taps each."""
    rng = np.random.default_rng(seed)
    widths = np.asarray(widths, np.int32)
    sizes = widths.astype(np.int64) ** 2 * os_ * os_
    c0 = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    wp = rng.integers(0, widths.size, nVis).astype(np.int32)
    s = widths[wp]
    iu = np.array([rng.integers(x // 2, gSize - x // 2) for x in s], np.int32)
    iv = np.array([rng.integers(0, gSize - x + 1) for x in s], np.int32)
    co = (c0[wp] + rng.integers(0, os_ * os_, nVis) * s.astype(np.int64) ** 2).astype(np.int32)
    n = int(sizes.sum())
    C = ((rng.integers(-4, 5, n) + 1j * rng.integers(-4, 5, n)) / 4).astype(np.complex64)
    data = (rng.integers(-3, 4, nVis) + 1j * rng.integers(-3, 4, nVis)).astype(np.complex64)
    return iu, iv, co, wp, data, C, widths


def test_noalloc_matches_grid_plain_twice_second_smaller(ctx):
    # first call: 20 000 visibilities of width 129 on 1024^2, exact inputs
    Sz, gSize = 129, 1024
    iu, iv, co, data, C, G = exact_problem(901, 20000, gSize, Sz, cluster=None)
    vis = tcv.Visibilities(ctx, iu, iv, co, [Sz], gSize, C.size)
    a, b = _both(ctx, vis, data, C, gSize, grid0=G)
    assert np.array_equal(a, b)
    assert np.abs(a - G).max() > 0          # something was gridded
    # second call, smaller: 1 500 visibilities of width 65 on 512^2, random complex
    Sz2, g2 = 65, 512
    iu, iv, co, data, C = small_problem(902, 1500, g2, Sz2)
    vis2 = tcv.Visibilities(ctx, iu, iv, co, [Sz2], g2, C.size)
    a2, b2 = _both(ctx, vis2, data, C, g2)
    assert rel_err(a2, b2) <= TOL


def test_noalloc_per_plane_widths(ctx):
    gSize = 700
    iu, iv, co, wp, data, C, widths = per_plane_problem(903, 6000, gSize)
    vis = tcv.Visibilities(ctx, iu, iv, co, widths, gSize, C.size, wPlane=wp)
    a, b = _both(ctx, vis, data, C, gSize)
    assert np.array_equal(a, b)
    # and again with fewer visibilities through the same array
    k = 900
    vis2 = tcv.Visibilities(ctx, iu[:k], iv[:k], co[:k], widths, gSize, C.size, wPlane=wp[:k])
    a2, b2 = _both(ctx, vis2, data[:k], C, gSize)
    assert np.array_equal(a2, b2)


def test_noalloc_refuses_jobs_beyond_the_scratch_array(ctx):
    """A plan larger than the array is refused on the host before anything is
    queued; the context stays usable."""
    Sz, gSize = 129, 4096
    n = 4_000_000                      # ~6 entries each: ~0.8 GB of plan
    d = ctx.empty(4, np.int32)         # never read: the call is refused first
    g = tcv._lib.Geometry()
    g.nVis, g.gSize, g.nPlanes, g.cSize = n, gSize, 1, Sz * Sz * 64 * 33
    g.d_iu = g.d_iv = g.d_cOffset = d.ptr
    g.d_wPlane = None
    sz = np.array([Sz], np.int32)
    g.sSize = sz.ctypes.data_as(ct.c_void_p)
    st = tcv.lib().tcv_grid_noalloc(ctx.h, ct.byref(g), d.ptr, d.ptr, d.ptr)
    assert st == 6, st                 # TCV_ERR_UNSUPPORTED
    assert b"scratch" in tcv.lib().tcv_last_error(ctx.h)
    ctx.sync()
    # empty and null-pointer calls
    g.nVis = 0
    assert tcv.lib().tcv_grid_noalloc(ctx.h, ct.byref(g), None, None, None) == 0
    g.nVis = 10
    assert tcv.lib().tcv_grid_noalloc(ctx.h, ct.byref(g), None, d.ptr, d.ptr) == 1
