"""GPU parity: libtcv's kernels against the oracle (CPU restatement of the
reference, itself bit-exact with the compiled reference) and the golden
fixtures.  Bars: index set-up bit-exact; grid and degridded values
max|d|/max|ref| <= 1e-5 on real and imaginary parts."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

import oracle as O
import tcv
from helpers import exact_problem, golden, grid_cols_choices, rel_err, small_problem
from tcv import setup as S

pytestmark = pytest.mark.gpu
TOL = 1e-5
TILES = [(64, 64), (64, 32), (32, 64), (32, 32), (16, 64), (16, 32), (8, 64), (8, 32)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def vis_from_host(ctx, iu, iv, co, sSize, gSize, cSize, wPlane=None):
    return tcv.Visibilities(ctx, iu, iv, co, sSize, gSize, cSize, wPlane=wPlane)


def gpu_grid(ctx, vis, data, C, gSize, method, tuning=None, grid0=None):
    d_data = ctx.to_device(data)
    d_C = ctx.to_device(C)
    g = ctx.to_device(grid0) if grid0 is not None else ctx.zeros((gSize, gSize), np.complex64)
    if method == "plain":
        tcv.grid_plain(ctx, vis, d_data, d_C, g)
    else:
        plan = tcv.Plan(ctx, vis, **(tuning or {}))
        plan.grid(d_data, d_C, g)
    return g.download()


def gpu_degrid(ctx, vis, grid, C, method, tuning=None):
    d_grid = ctx.to_device(grid)
    d_C = ctx.to_device(C)
    out = ctx.empty(vis.nVis, np.complex64)
    if method == "plain":
        tcv.degrid_plain(ctx, vis, d_grid, d_C, out)
    else:
        plan = tcv.Plan(ctx, vis, **(tuning or {}))
        plan.degrid(d_grid, d_C, out)
    return out.download()


# ---- a6: index set-up, bit-exact ---------------------------------------------

def test_offsets_attic_bitexact_vs_reference(ctx):
    meta, arrs = golden("attic")
    u, v, w = S.attic_uvw(meta["nSamples"], meta["baseline"])
    freq = S.attic_freq(1)
    sup, wcs, _ = S.attic_initC(freq[0], meta["cellSize"], meta["wSize"], meta["baseline"], 8,
                                table=False)
    assert sup == meta["support"] and wcs == meta["wCellSize"]
    iu, iv, wp, co, npp, bad = tcv.init_offsets(
        ctx, ctx.to_device(u), ctx.to_device(v), ctx.to_device(w), freq, meta["cellSize"], wcs,
        meta["wSize"], meta["gSize"], 8, [2 * sup + 1], None)
    assert bad == 0
    assert sha(iu.download()) == meta["sha256"]["iu"]
    assert sha(iv.download()) == meta["sha256"]["iv"]
    assert sha(co.download()) == meta["sha256"]["cOffset"]
    # w-plane histogram: only planes 8..24 are hit by [-1000, 1000] m
    assert npp.sum() == meta["nSamples"] and npp[:8].sum() == 0 and npp[25:].sum() == 0


@pytest.mark.parametrize("rt", [0, 1, 2, 3, 4])
def test_offsets_current_bitexact_vs_reference(ctx, rt):
    meta, arrs = golden(f"current_rt{rt}")
    prob = S.current_problem(rt)
    assert sha(prob.C) == meta["sha256"]["C"]
    iu, iv, wp, co, npp, bad = tcv.init_offsets(
        ctx, ctx.to_device(prob.u), ctx.to_device(prob.v), ctx.to_device(prob.w), prob.freq,
        prob.cellSize, prob.wCellSize, prob.wSize, prob.gSize, prob.overSample, prob.sSize,
        prob.cOffset0)
    assert sha(iu.download()) == meta["sha256"]["iu"]
    assert sha(iv.download()) == meta["sha256"]["iv"]
    assert sha(wp.download()) == meta["sha256"]["wPlane"]
    assert sha(co.download()) == meta["sha256"]["cOffset"]
    assert np.array_equal(npp, arrs["numPerPlane"])


def test_offsets_spectral_vs_oracle(ctx):
    prob = S.spectral_problem(nSamples=1 << 16)
    iu, iv, wp, co, npp, bad = tcv.init_offsets(
        ctx, ctx.to_device(prob.u), ctx.to_device(prob.v), ctx.to_device(prob.w), prob.freq,
        prob.cellSize, prob.wCellSize, prob.wSize, prob.gSize, 8, prob.sSize, None)
    ref = O.attic_initCOffset(prob.u, prob.v, prob.w, prob.freq, prob.cellSize, prob.wCellSize,
                              prob.wSize, prob.gSize, prob.support, 8)
    assert bad == 0
    for a, b in zip((iu, iv, co), ref):
        assert np.array_equal(a.download(), b)


# ---- a7/a8 on small seeded problems --------------------------------------------

@pytest.mark.parametrize("method", ["plain", "tiled"])
def test_single_visibility_patch_is_C_slice(ctx, method):
    rng = np.random.default_rng(5)
    Sz, gSize = 129, 512
    iu, iv, co, data, C = small_problem(5, 1, gSize, Sz)
    data[:] = 1.0 + 0.0j
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    g = gpu_grid(ctx, vis, data, C, gSize, method)
    u0, v0 = iu[0] - Sz // 2, iv[0]
    patch = g[v0:v0 + Sz, u0:u0 + Sz]
    assert np.array_equal(patch, C[co[0]:co[0] + Sz * Sz].reshape(Sz, Sz))
    g[v0:v0 + Sz, u0:u0 + Sz] = 0
    assert not g.any()
    del rng


@pytest.mark.parametrize("Sz", [129, 65, 33, 7])
@pytest.mark.parametrize("method", ["plain", "tiled"])
def test_grid_degrid_small_vs_oracle(ctx, method, Sz):
    gSize = 384
    iu, iv, co, data, C = small_problem(11 + Sz, 3000, gSize, Sz, complex_=True)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    g = gpu_grid(ctx, vis, data, C, gSize, method)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    assert rel_err(g, gref) <= TOL
    g64 = O.grid_f64(data, iu, iv, co, C, gSize, Sz)
    assert rel_err(g, g64) <= TOL
    # degrid the oracle's grid
    out = gpu_degrid(ctx, vis, gref, C, method)
    oref = O.degrid(gref, iu, iv, co, C, Sz)
    assert rel_err(out, oref) <= TOL


@pytest.mark.parametrize("tw,th", TILES)
def test_tiled_shapes_vs_oracle(ctx, tw, th):
    Sz, gSize = 93, 300
    iu, iv, co, data, C = small_problem(100 + tw + th, 1500, gSize, Sz, cluster=60)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    for tuning in (dict(tile_w=tw, tile_h=th),
                   dict(tile_w=tw, tile_h=th, slice_keys=1, phases=1),
                   dict(tile_w=tw, tile_h=th, chunk=64, slice_keys=256, phases=4),
                   dict(tile_w=tw, tile_h=th, chunk=96, slice_keys=320, phases=16, xcd_split=1),
                   dict(tile_w=tw, tile_h=th, kernel=2, degrid_kernel=2),
                   dict(tile_w=tw, tile_h=th, chunk=96, slice_keys=320, phases=16, xcd_split=1,
                        kernel=2, degrid_kernel=1),
                   dict(tile_w=tw, tile_h=th, kernel=3, degrid_kernel=3),
                   dict(tile_w=tw, tile_h=th, chunk=96, slice_keys=320, phases=16, xcd_split=1,
                        kernel=3, degrid_kernel=3),
                   dict(tile_w=tw, tile_h=th, chunk=96, slice_keys=320, phases=16, grid_phases=4,
                        xcd_split=1, kernel=3, degrid_kernel=3)):
        g = gpu_grid(ctx, vis, data, C, gSize, "tiled", tuning)
        assert rel_err(g, gref) <= TOL, tuning
        out = gpu_degrid(ctx, vis, gref, C, "tiled", tuning)
        assert rel_err(out, O.degrid(gref, iu, iv, co, C, Sz)) <= TOL, tuning


@pytest.mark.parametrize("tw,th", [(128, 64), (64, 128), (32, 256)])
@pytest.mark.parametrize("Sz,gSize", [(129, 300), (97, 257), (5, 200)])
def test_wide_tiles_pair_kernels_vs_oracle(ctx, Sz, gSize, tw, th):
    """Tiles wider or taller than 64 (pair kernels only; 128 x 64 is the
    default for kernels 89 wide and up): grid and degrid within 1e-5 of the
    oracle over item splits, phase counts and the XCD queues, on grids that are
    not a multiple of the tile (257, 300) and with footprints narrower than the
    tile."""
    iu, iv, co, data, C = small_problem(160 + Sz, 1500, gSize, Sz, cluster=60)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    oref = O.degrid(gref, iu, iv, co, C, Sz)
    for tuning in (dict(tile_w=tw, tile_h=th),
                   dict(tile_w=tw, tile_h=th, slice_keys=1, phases=1),
                   dict(tile_w=tw, tile_h=th, chunk=64, slice_keys=256, phases=4),
                   dict(tile_w=tw, tile_h=th, chunk=96, slice_keys=320, phases=16, xcd_split=1),
                   dict(tile_w=tw, tile_h=th, chunk=96, slice_keys=320, phases=16, grid_phases=4,
                        xcd_split=1)):
        info = tcv.Plan(ctx, vis, **tuning).info()
        assert (info["tile_w"], info["tile_h"], info["kernel"], info["degrid_kernel"]) == (tw, th, 3, 3)
        out = gpu_degrid(ctx, vis, gref, C, "tiled", tuning)
        assert rel_err(out, oref) <= TOL, tuning
        # the gridder on every column-group width (one phase and one key: each
        # tile's work item is its only one, flushed by plain read-modify-write)
        for cols in grid_cols_choices(tw, th):
            g = gpu_grid(ctx, vis, data, C, gSize, "tiled", dict(tuning, grid_cols=cols))
            assert rel_err(g, gref) <= TOL, (tuning, cols)


def test_wide_tiles_default_and_rejections(ctx):
    """Kernels 89 wide and up get 128 x 64 tiles by default (the pair gridder
    on 64-column groups), narrower ones 64-wide tiles; 128-wide tiles with a
    one-column kernel (1, 2) or with width-1 kernels, and column groups that
    are not built, are refused with an error."""
    iu, iv, co, data, C = small_problem(170, 200, 400, 129)
    vis = vis_from_host(ctx, iu, iv, co, [129], 400, C.size)
    info = tcv.Plan(ctx, vis).info()
    assert (info["tile_w"], info["tile_h"], info["grid_cols"]) == (128, 64, 64)
    assert tcv.Plan(ctx, vis, kernel=1).info()["tile_w"] == 64
    assert tcv.Plan(ctx, vis, tile_w=64, tile_h=64).info()["grid_cols"] == 64
    assert tcv.Plan(ctx, vis, tile_w=128, tile_h=64, grid_cols=32).info()["grid_cols"] == 32
    for bad in (dict(tile_w=128, tile_h=64, kernel=1), dict(tile_w=128, tile_h=64, degrid_kernel=2),
                dict(tile_w=64, tile_h=128, kernel=2), dict(tile_w=32, tile_h=256, degrid_kernel=1),
                dict(tile_w=128, tile_h=64, grid_cols=48), dict(tile_w=128, tile_h=64, grid_cols=16),
                dict(tile_w=32, tile_h=256, grid_cols=16), dict(tile_w=64, tile_h=32, grid_cols=8),
                dict(tile_w=128, tile_h=64, grid_cols=-1)):
        with pytest.raises(tcv.TcvError):
            tcv.Plan(ctx, vis, **bad)
    iu, iv, co, data, C = small_problem(171, 200, 256, 89)
    vis = vis_from_host(ctx, iu, iv, co, [89], 256, C.size)
    assert tcv.Plan(ctx, vis).info()["tile_w"] == 128
    iu, iv, co, data, C = small_problem(173, 200, 256, 87)
    vis = vis_from_host(ctx, iu, iv, co, [87], 256, C.size)
    assert tcv.Plan(ctx, vis).info()["tile_w"] == 64
    iu, iv, co, data, C = small_problem(172, 200, 64, 1)
    vis = vis_from_host(ctx, iu, iv, co, [1], 64, C.size)
    with pytest.raises(tcv.TcvError):
        tcv.Plan(ctx, vis, tile_w=128, tile_h=64)


def test_grid_accumulates_into_existing_grid(ctx):
    Sz, gSize = 65, 256
    iu, iv, co, data, C = small_problem(21, 500, gSize, Sz)
    rng = np.random.default_rng(3)
    g0 = (rng.standard_normal((gSize, gSize)) + 1j * rng.standard_normal((gSize, gSize))).astype(np.complex64)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    ref = O.grid(data, iu, iv, co, C, gSize, Sz, out=g0.copy())
    for method in ("plain", "tiled"):
        assert rel_err(gpu_grid(ctx, vis, data, C, gSize, method, grid0=g0), ref) <= TOL


def test_adjointness_fp64(ctx):
    """<G, grid(d)> == <degrid(G), d> (bilinear, no conjugation), in FP64."""
    Sz, gSize = 129, 400
    iu, iv, co, data, C = small_problem(31, 800, gSize, Sz)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    rng = np.random.default_rng(4)
    G = (rng.standard_normal((gSize, gSize)) + 1j * rng.standard_normal((gSize, gSize))).astype(np.complex64)
    for method in ("plain", "tiled"):
        gd = gpu_grid(ctx, vis, data, C, gSize, method).astype(np.complex128)
        dg = gpu_degrid(ctx, vis, G, C, method).astype(np.complex128)
        lhs = np.sum(G.astype(np.complex128) * gd)
        rhs = np.sum(dg * data.astype(np.complex128))
        assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("Sz", [129, 33])
def test_fp64_reference_kernels_vs_oracle(ctx, Sz):
    """The device's FP64 reference mode equals the oracle's FP64 gridding and
    degridding (same float-stored taps and data, FP64 products and sums) up to
    FP64 summation order; PSF mode grids data = 1 + 0i."""
    gSize, n = 384, 2000
    iu, iv, co, data, C = small_problem(41 + Sz, n, gSize, Sz)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    d_data, d_C = ctx.to_device(data), ctx.to_device(C)
    g = ctx.zeros((gSize, gSize), np.complex128)
    tcv.grid_plain_f64(ctx, vis, d_data, d_C, g)
    ref = O.grid_f64(data, iu, iv, co, C, gSize, Sz)
    assert rel_err(g.download(), ref) <= 1e-12
    # a dropped cross term (im = dr*ci only) would be off by O(1)
    assert rel_err(g.download(), O.grid(data, iu, iv, co, C, gSize, Sz)) <= TOL
    gp = ctx.zeros((gSize, gSize), np.complex128)
    tcv.grid_plain_f64(ctx, vis, None, d_C, gp)
    ones = np.ones(n, np.complex64)
    assert rel_err(gp.download(), O.grid_f64(ones, iu, iv, co, C, gSize, Sz)) <= 1e-12
    rng = np.random.default_rng(Sz)
    G = (rng.standard_normal((gSize, gSize)) + 1j * rng.standard_normal((gSize, gSize))).astype(np.complex64)
    out = ctx.empty(n, np.complex128)
    tcv.degrid_plain_f64(ctx, vis, ctx.to_device(G), d_C, out)
    assert rel_err(out.download(), O.degrid_f64(G, iu, iv, co, C, Sz)) <= 1e-12


def test_out_of_range_visibilities_are_skipped(ctx):
    Sz, gSize = 33, 128
    iu, iv, co, data, C = small_problem(41, 200, gSize, Sz)
    iu[:5] = 3            # footprint crosses the left edge
    iv[5:10] = gSize - 2  # crosses the bottom edge
    co[10:12] = C.size    # slice past the table
    good = np.ones(iu.size, bool)
    good[:12] = False
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    ref = O.grid(data[good], iu[good], iv[good], co[good], C, gSize, Sz)
    g = gpu_grid(ctx, vis, data, C, gSize, "plain")
    assert ctx.range_errors() == 12
    assert rel_err(g, ref) <= TOL
    plan = tcv.Plan(ctx, vis)
    assert plan.info()["out_of_range"] == 12
    g2 = ctx.zeros((gSize, gSize), np.complex64)
    plan.grid(ctx.to_device(data), ctx.to_device(C), g2)
    assert rel_err(g2.download(), ref) <= TOL
    out = ctx.empty(iu.size, np.complex64)
    plan.degrid(ctx.to_device(ref), ctx.to_device(C), out)
    o = out.download()
    assert not o[~good].any()
    assert rel_err(o[good], O.degrid(ref, iu[good], iv[good], co[good], C, Sz)) <= TOL


# ---- the drop-in Benchmark surfaces ---------------------------------------------

@pytest.mark.parametrize("method", ["tiled", "plain"])
def test_attic_benchmark_vs_reference(ctx, method):
    """Full config-1 workload through the drop-in Benchmark; checked against the
    compiled reference's fixture pixels and the oracle's full grid/degrid."""
    meta, arrs = golden("attic")
    b = tcv.Benchmark(ctx, method=method)
    b.init()
    assert b.getSupport() == 64 and b.nOutOfRange == 0
    iu, iv, co = b.samples_host()
    assert sha(co) == meta["sha256"]["cOffset"]
    b.runGrid()
    g = b.grid_host()
    gv = g.reshape(-1)[arrs["grid_pix_idx"]]
    assert rel_err(gv, arrs["grid_pix_val"]) <= TOL
    C = b.C.download()
    gref = O.grid(np.ones(iu.size, np.complex64), iu, iv, co, C, 4096, 129)
    assert rel_err(g, gref) <= TOL
    # degrid the reference's own grid so both sides see identical input
    b.grid.upload(gref)
    b.runDegrid()
    out = b.outdata_host()
    assert rel_err(out[:256], arrs["outdata_head"]) <= TOL
    assert rel_err(out, O.degrid(gref, iu, iv, co, C, 129)) <= TOL


@pytest.mark.parametrize("rt", [0, 1, 3, 2])
def test_current_benchmark_vs_oracle(ctx, rt):
    """Variable per-plane widths and complex Fresnel C (runType 0, the continuum
    case: 483 planes on a 6440^2 grid; runType 1, spectral), small kernels with
    overSample 128 (runType 3), nearest neighbour (runType 2).

    These ASKAP-layout workloads pile millions of visibilities onto the central
    pixels, where the reference's own sequential FP32 sum is off the exact sum
    by up to 3.5e-5 of the maximum (measured against the FP64 oracle).  So the
    1e-5 bar is applied against the FP64 oracle, and against the FP32
    reference the GPU must be within the reference's own error plus 1e-5."""
    meta, arrs = golden(f"current_rt{rt}")
    b = tcv.CurrentBenchmark(ctx)
    b.setRunType(rt)
    b.init()
    assert b.nPixelsGridded() == meta["nPixelsGridded"]
    b.runGrid()
    g = b.grid_host()
    iu, iv = b.samples.iu.download(), b.samples.iv.download()
    wp, co = b.samples.wPlane.download(), b.samples.cOffset.download()
    C = b.C.download()
    ones = np.ones(iu.size, np.complex64)
    gref = O.grid(ones, iu, iv, co, C, b.gSize, b.sSize, wPlane=wp)
    g64 = O.grid_f64(ones, iu, iv, co, C, b.gSize, b.sSize, wPlane=wp)
    assert rel_err(g, g64) <= TOL
    ref_err = rel_err(gref, g64)
    assert rel_err(g, gref) <= ref_err + TOL
    # the compiled reference's own pixels (bit-identical to gref)
    idx = arrs["grid_pix_idx"]
    gv = g.reshape(-1)[idx]
    assert rel_err(gv, arrs["grid_pix_val"]) <= rel_err(arrs["grid_pix_val"], g64.reshape(-1)[idx]) + TOL
    # degrid the reference's grid so both sides see the same input
    b.grid1.upload(gref)
    b.runDegrid()
    out = b.outdata_host()
    oref = O.degrid(gref, iu, iv, co, C, b.sSize, wPlane=wp)
    assert rel_err(out, oref) <= TOL
    assert rel_err(out[:256], arrs["outdata_head"]) <= TOL


def test_current_benchmark_rt4_large_kernels(ctx):
    """runType 4: 87 x 87 kernels, 5.4 M visibilities on a 4129^2 grid (4.1e10
    grid points, too many for a full oracle run in a test).  Every cell of the
    tiled grid against the device's FP64 reference kernels, windows of the
    densest centre, a tile-straddling spot and the populated rim against the
    oracle (FP32 and FP64), and sampled degridded values against the oracle."""
    b = tcv.CurrentBenchmark(ctx)
    b.setRunType(4)
    b.init()
    meta, _ = golden("current_rt4")
    assert b.gSize == meta["gSize"] and int(b.sSize[0]) == 87
    assert b.nPixelsGridded() == meta["nPixelsGridded"] and b.nOutOfRange == 0
    b.runGrid()
    g = b.grid_host()
    g64d = ctx.zeros((b.gSize, b.gSize), np.complex128)
    tcv.grid_plain_f64(ctx, b.samples, b.data, b.C, g64d)
    g64 = g64d.download()
    g64d.free()
    e = rel_err(g, g64)
    print(f"runType 4 grid: max error vs FP64 {e:.2e}")
    assert e <= TOL
    iu, iv = b.samples.iu.download(), b.samples.iv.download()
    co = b.samples.cOffset.download()
    C = b.C.download()
    ones = np.ones(iu.size, np.complex64)
    c = b.gSize // 2
    ys, xs = np.nonzero(np.abs(g64) > 0)
    k = int(np.argmax(xs))             # the populated cell furthest right
    rim = (min(int(xs[k]) - 100, b.gSize - 128), min(max(int(ys[k]) - 32, 0), b.gSize - 64))
    for x0, y0 in ((c - 64, c - 32), (c + 301, c - 517), rim):
        w32 = O.grid_window(ones, iu, iv, co, C, b.gSize, 87, x0, y0, 128, 64)
        w64 = O.grid_window(ones, iu, iv, co, C, b.gSize, 87, x0, y0, 128, 64, f64=True)
        assert w64.any(), (x0, y0)
        assert rel_err(g[y0:y0 + 64, x0:x0 + 128], w64) <= TOL, (x0, y0)
        assert rel_err(g64[y0:y0 + 64, x0:x0 + 128], w64) <= 1e-12, (x0, y0)
        del w32
    b.runDegrid()
    out = b.outdata_host()
    pick = np.sort(np.random.default_rng(9).choice(iu.size, 3000, replace=False))
    want = O.degrid(g, iu[pick], iv[pick], co[pick], C, 87)
    assert rel_err(out[pick], want) <= TOL


def test_current_benchmark_setsort(ctx):
    """setSort(1) (current/tConvolveMPI/Benchmark.cc:275-298): the index arrays
    are the reference's counting sort by w-plane of the unsorted arrays, and
    the grid (data = 1 everywhere, so order-free) is unchanged within 1e-5."""
    rt = 1
    base = tcv.CurrentBenchmark(ctx)
    base.setRunType(rt)
    base.init()
    srt = tcv.CurrentBenchmark(ctx)
    srt.setRunType(rt)
    srt.setSort(1)
    srt.init()
    h0 = [a.download() for a in (base.samples.iu, base.samples.iv, base.samples.wPlane,
                                 base.samples.cOffset)]
    want = O.sort_by_plane(*h0, base.numPerPlane)
    got = [a.download() for a in (srt.samples.iu, srt.samples.iv, srt.samples.wPlane,
                                  srt.samples.cOffset)]
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert srt.nPixelsGridded() == base.nPixelsGridded()
    base.runGrid()
    srt.runGrid()
    assert rel_err(srt.grid_host(), base.grid_host()) <= TOL


def test_attic_wide_vs_oracle(ctx):
    """Config 1 as BASELINE states it: u, v, w uniform in [-2000, 2000] m (the
    reference generator with a 4000 m span), every one of the 33 w-planes in
    use; the tiled grid and degrid against the oracle's full runs at 1e-5."""
    prob = S.attic_problem("wide")
    job = tcv.GridJob(ctx, prob)
    assert job.nBad == 0
    assert (job.numPerPlane > 0).all() and job.numPerPlane.size == 33
    iu, iv, _, co = job.offsets_host()
    job.zero_grid()
    job.grid("tiled")
    g = job.grid_d.download()
    gref = O.grid(prob.data, iu, iv, co, prob.C, 4096, 129)
    assert rel_err(g, gref) <= TOL
    assert rel_err(g, O.grid_f64(prob.data, iu, iv, co, prob.C, 4096, 129)) <= TOL
    job.grid_d.upload(gref)
    job.degrid("tiled")
    out = job.out.download()
    assert rel_err(out, O.degrid(gref, iu, iv, co, prob.C, 129)) <= TOL


def test_cpp_dropin_driver(ctx):
    """The tConvolveMPI-style C++ driver over the header-only drop-in
    (include/tcv_benchmark.hpp) runs the attic workload on the device; its grid and
    degrid checksums match the compiled reference's."""
    import os
    import subprocess
    meta, _ = golden("attic")
    exe = os.path.join(os.path.dirname(tcv.LIB_PATH), "tconvolve_gpu")
    for flag in ([], ["--plain"]):
        r = subprocess.run([exe, "--checksum"] + flag, capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stderr
        assert "Gridding rate" in r.stdout and "Degridding rate" in r.stdout
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("checksum")][0].split()
        gr, gi = float(line[2]), float(line[3])
        assert gi == 0.0
        assert abs(gr - meta["grid_sum"][0]) <= 1e-5 * abs(meta["grid_sum"][0])
        assert line[-1] == "0"


def _driver_checksums(out):
    rows = {}
    for ln in out.splitlines():
        if ln.startswith("checksum run"):
            f = ln.split()
            rows[int(f[2])] = dict(pixels=int(f[4]), grid=complex(float(f[6]), float(f[7])),
                                   outdata=complex(float(f[9]), float(f[10])), bad=int(f[12]))
    return rows


def test_cpp_current_driver(ctx):
    """The current tConvolveMPI driver (C++, over the drop-in current Benchmark.h)
    on runType 1: the gridded-pixel count equals the compiled reference's; the
    grid's sum equals the FP64 oracle grid's (data = 1: sum_v sum(slice_v))
    within 2e-6 of the sum of its magnitudes; the degridded data's sum equals
    sum_p G[p]^2 (the degridding of the gridded grid, bilinear); with setSort(1)
    the sums are unchanged.  The reference's two-run loop (runType 0 then 1 on
    one instance) completes and reports both."""
    import os
    import subprocess
    meta, _ = golden("current_rt1")
    exe = os.path.join(os.path.dirname(tcv.LIB_PATH), "tconvolve_current_gpu")
    p = O.current_params(1)
    u, v, w = O.current_uvw(p["baseline"])
    sS, c0, C = O.current_initC(p["uvCellSize"], p["wSize"], p["support"], p["overSample"],
                                p["wCellSize"])
    iu, iv, wp, co, npp = O.current_initCOffset(u, v, w, np.array([1.0e9 / 2.998e8]), p["uvCellSize"],
                                                p["wCellSize"], p["wSize"], p["gSize"], p["overSample"],
                                                sS, c0)
    g64 = O.grid_f64(np.ones(iu.size, np.complex64), iu, iv, co, C, p["gSize"], sS, wPlane=wp)
    sums = {}
    for flags in ([], ["--sort"]):
        r = subprocess.run([exe, "--runtypes", "1", "--checksum"] + flags, capture_output=True, text=True,
                           timeout=300)
        print(r.stdout[-2000:])
        assert r.returncode == 0, r.stderr
        assert "Spectral gridding performance" in r.stdout and "Spectral degridding performance" in r.stdout
        row = _driver_checksums(r.stdout)[1]
        assert row["pixels"] == meta["nPixelsGridded"] and row["bad"] == 0
        assert abs(row["grid"] - g64.sum()) <= 2e-6 * np.abs(g64).sum(), (row["grid"], g64.sum())
        want = (g64 * g64).sum()
        assert abs(row["outdata"] - want) <= 1e-5 * (np.abs(g64) ** 2).sum(), (row["outdata"], want)
        sums[tuple(flags)] = row
    r = subprocess.run([exe, "--checksum"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Done") == 2 and "continuum gridding rate (per process)" in r.stdout
    rows = _driver_checksums(r.stdout)
    assert set(rows) == {0, 1} and all(x["bad"] == 0 for x in rows.values())
    assert rows[0]["pixels"] == golden("current_rt0")[0]["nPixelsGridded"]


def test_cuda_gridder_replacement(ctx, tmp_path):
    """The CudaGridder.cc replacement (gridKernelCuda / degridKernelCuda with the
    signatures of attic/tConvolveCuda/CudaGridder.h:32-50, declared in the
    calling program itself) against the oracle: gridding accumulates into the
    given grid, degridding overwrites the data."""
    import os
    import subprocess
    Sz, gSize = 129, 512
    iu, iv, co, data, C = small_problem(71, 3000, gSize, Sz)
    rng = np.random.default_rng(72)
    grid0 = (rng.standard_normal((gSize, gSize)) + 1j * rng.standard_normal((gSize, gSize))).astype(np.complex64)
    gridin = (rng.standard_normal((gSize, gSize)) + 1j * rng.standard_normal((gSize, gSize))).astype(np.complex64)
    for name, a in (("iu.i32", iu), ("iv.i32", iv), ("cOffset.i32", co), ("data.c64", data), ("C.c64", C),
                    ("grid0.c64", grid0), ("gridin.c64", gridin)):
        a.tofile(str(tmp_path / name))
    (tmp_path / "params.txt").write_text(f"{gSize} {Sz // 2}\n")
    exe = os.path.join(os.path.dirname(tcv.LIB_PATH), "cuda_gridder_check")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "grid_s" in r.stdout and not r.stderr
    g = np.fromfile(str(tmp_path / "grid.c64"), np.complex64).reshape(gSize, gSize)
    d = np.fromfile(str(tmp_path / "degrid.c64"), np.complex64)
    assert rel_err(g, O.grid(data, iu, iv, co, C, gSize, Sz, out=grid0.copy())) <= TOL
    assert rel_err(d, O.degrid(gridin, iu, iv, co, C, Sz)) <= TOL


def test_profiling_counts_launches(ctx):
    Sz, gSize = 33, 128
    iu, iv, co, data, C = small_problem(51, 100, gSize, Sz)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    ctx.set_profiling(True)
    ctx.reset_stats()
    gpu_grid(ctx, vis, data, C, gSize, "tiled")
    n, ms = ctx.kernel_stats("grid_tiled")
    ctx.set_profiling(False)
    assert n == 1 and ms > 0


def test_copy_kernel_bit_exact(ctx):
    rng = np.random.default_rng(5)
    for n in (0, 4, 1 << 10, (1 << 20) + 12):
        a = rng.integers(0, 2**32, n, dtype=np.uint32)
        src = ctx.to_device(a)
        dst = ctx.empty(n, np.uint32)
        dst.copy_from(src)
        assert np.array_equal(dst.download(), a)


def test_read_bandwidth_probe(ctx):
    """The streaming-read probe returns a positive rate for a power-of-two
    buffer and refuses other sizes before launching anything."""
    buf = ctx.empty((4 << 20) // 4, np.uint32)
    buf.zero()
    g = ctx.read_bandwidth(buf, rounds=64)
    assert 100.0 < g < 1e6
    with pytest.raises(tcv.TcvError):
        ctx.read_bandwidth(ctx.empty(3000, np.uint32))
    with pytest.raises(tcv.TcvError):
        ctx.read_bandwidth(buf, rounds=12)


@pytest.mark.parametrize("tw,th", [(64, 64), (32, 64), (16, 32)])
def test_lds_and_register_kernels_bitwise_on_exact_inputs(ctx, tw, th):
    """On order-independent (exactly representable) sums the register-tile and
    LDS-tile kernels, with split or whole work items, equal the oracle bit for
    bit: every visibility reaches every cell of its footprint exactly once."""
    Sz, gSize = 129, 400
    iu, iv, co, data, C, G = exact_problem(77 + tw, 2500, gSize, Sz, cluster=100)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    oref = O.degrid(G, iu, iv, co, C, Sz)
    for kernel in (1, 2, 3):
        for t in (dict(chunk=1 << 20, slice_keys=1, phases=1),
                  dict(chunk=64, slice_keys=64, phases=8, xcd_split=1),
                  dict(chunk=64, slice_keys=64, phases=16, grid_phases=3, xcd_split=1)):
            tuning = dict(t, tile_w=tw, tile_h=th, kernel=kernel, degrid_kernel=kernel)
            assert np.array_equal(gpu_grid(ctx, vis, data, C, gSize, "tiled", tuning), gref), tuning
            assert np.array_equal(gpu_degrid(ctx, vis, G, C, "tiled", tuning), oref), tuning


@pytest.mark.parametrize("Sz", [3, 5, 7, 33, 65, 129])
def test_pair_kernels_every_width_exact(ctx, Sz):
    """The two-tap kernels at every footprint parity and width, including
    footprints that start or end on a tile edge: bit-exact on exact inputs."""
    gSize = 256
    iu, iv, co, data, C, G = exact_problem(300 + Sz, 1500, gSize, Sz, cluster=60)
    # add footprints at both grid edges and at every offset around a tile edge
    half = Sz // 2
    extra_u = np.array([half, gSize - half - 1] + [64 + d for d in range(-Sz, 2)], np.int32)
    extra_u = np.clip(extra_u, half, gSize - half - 1)
    extra_v = np.clip(np.arange(extra_u.size, dtype=np.int32) * 7 % (gSize - Sz + 1), 0, gSize - Sz)
    iu = np.concatenate([iu, extra_u])
    iv = np.concatenate([iv, extra_v.astype(np.int32)])
    co = np.concatenate([co, co[:extra_u.size]])
    rng = np.random.default_rng(Sz)
    data = np.concatenate([data, (rng.integers(-3, 4, extra_u.size) + 0j).astype(np.complex64)])
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    oref = O.degrid(G, iu, iv, co, C, Sz)
    for tw, th in ((128, 64), (64, 128), (32, 256), (64, 64), (64, 32), (32, 64), (16, 32), (8, 64)):
        tuning = dict(tile_w=tw, tile_h=th, kernel=3, degrid_kernel=3, phases=4, slice_keys=16, chunk=32)
        assert np.array_equal(gpu_degrid(ctx, vis, G, C, "tiled", tuning), oref), (tw, th)
        # every column-group width of the pair gridder
        for cols in grid_cols_choices(tw, th):
            assert np.array_equal(gpu_grid(ctx, vis, data, C, gSize, "tiled", dict(tuning, grid_cols=cols)),
                                  gref), (tw, th, cols)


@pytest.mark.parametrize("Sz,keys", [(3, 4), (9, 4), (33, 4), (129, 4), (9, 4096), (33, 1)])
def test_pair_kernels_first_and_last_slice_exact(ctx, Sz, keys):
    """Footprints whose C slice is the table's first or last, at every offset
    around a tile edge: the pair kernels then load the pair that starts or ends
    on the edge tap (no tap outside the table is read); bit-exact.  The
    gridder chooses that path per work item from its first and last entries'
    C-slice keys: also with more keys than slices (4096) and with one key."""
    gSize = 256
    iu, iv, co, data, C, G = exact_problem(500 + Sz, 400, gSize, Sz, cluster=None)
    half = Sz // 2
    n_slices = C.size // (Sz * Sz)
    extra_u = np.clip(np.array([half, gSize - half - 1] + [64 + d for d in range(-Sz - 1, 3)] +
                               [128 + d for d in range(-Sz - 1, 3)], np.int32), half, gSize - half - 1)
    extra_v = np.clip(np.arange(extra_u.size, dtype=np.int32) * 11 % (gSize - Sz + 1), 0, gSize - Sz)
    iu = np.concatenate([iu, extra_u, extra_u])
    iv = np.concatenate([iv, extra_v, (gSize - Sz - extra_v)]).astype(np.int32)
    first_last = np.array([0, (n_slices - 1) * Sz * Sz], np.int32)
    co = np.concatenate([co, first_last[np.arange(2 * extra_u.size) % 2]]).astype(np.int32)
    co[: co.size // 4] = first_last[np.arange(co.size // 4) % 2]
    rng = np.random.default_rng(Sz + 7)
    data = np.concatenate([data, (rng.integers(-3, 4, 2 * extra_u.size) +
                                  1j * rng.integers(-3, 4, 2 * extra_u.size)).astype(np.complex64)])
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    oref = O.degrid(G, iu, iv, co, C, Sz)
    for tw, th in ((128, 64), (64, 128), (32, 256), (64, 64), (32, 64), (16, 32)):
        tuning = dict(tile_w=tw, tile_h=th, kernel=3, degrid_kernel=3, phases=min(2, keys),
                      slice_keys=keys, chunk=48)
        assert np.array_equal(gpu_degrid(ctx, vis, G, C, "tiled", tuning), oref), (tw, th)
        for cols in grid_cols_choices(tw, th):
            assert np.array_equal(gpu_grid(ctx, vis, data, C, gSize, "tiled", dict(tuning, grid_cols=cols)),
                                  gref), (tw, th, cols)


@pytest.mark.parametrize("band", [(0.40, 0.45), (0.0, 0.02), (0.97, 1.0), (0.5, 0.5)])
def test_phases_over_the_keys_in_use_exact(ctx, band):
    """Phases split the slice keys the entries use: entries whose C slices lie
    in one band of the table (the middle, the start, the end, one slice) are
    cut into 1 to 64 phases, more phases than used keys included, over the
    three binnings (two-level with 16 keys a tile and with 4096, more keys than
    slices; one-level (tile, key) with global atomics on 8192 tiles); every
    entry lands in exactly one work item: bit-exact."""
    Sz = 9
    for keys, tw, th, gSize in ((16, 64, 64, 256), (4096, 128, 64, 256), (8, 16, 32, 2048)):
        iu, iv, co, data, C, G = exact_problem(1200 + int(100 * band[0]) + keys, 2000, gSize, Sz,
                                               cluster=None)
        n_slices = C.size // (Sz * Sz)
        lo = min(int(band[0] * n_slices), n_slices - 1)
        hi = max(lo, min(int(band[1] * n_slices), n_slices - 1))
        rng = np.random.default_rng(int(1000 * band[1]) + keys)
        co = (rng.integers(lo, hi + 1, iu.size) * Sz * Sz).astype(np.int32)
        vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
        gref = O.grid(data, iu, iv, co, C, gSize, Sz)
        oref = O.degrid(G, iu, iv, co, C, Sz)
        for phases in (1, 3, 8, 64):
            tuning = dict(tile_w=tw, tile_h=th, slice_keys=keys, phases=min(phases, keys),
                          grid_phases=min(phases, keys), chunk=64)
            assert np.array_equal(gpu_grid(ctx, vis, data, C, gSize, "tiled", tuning), gref), tuning
            assert np.array_equal(gpu_degrid(ctx, vis, G, C, "tiled", tuning), oref), tuning


@pytest.mark.parametrize("tw,th,keys", [(16, 32, 4), (64, 64, 1), (64, 64, 16)])
def test_plan_binning_paths_exact(ctx, tw, th, keys):
    """The three plan binnings give the same (bit-exact) results: (tile, key)
    with global atomics when there are more tiles than the LDS histogram holds
    (2048^2 grid in 16x32 tiles: 8192 tiles), (tile) only in LDS when there is
    one key, and the two-level binning (tiles in LDS, then keys per tile)."""
    Sz, gSize = 33, 2048
    iu, iv, co, data, C, G = exact_problem(900 + tw + keys, 3000, gSize, Sz, cluster=None)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    oref = O.degrid(G, iu, iv, co, C, Sz)
    tuning = dict(tile_w=tw, tile_h=th, slice_keys=keys, phases=min(keys, 4), chunk=64)
    info = tcv.Plan(ctx, vis, **tuning).info()
    assert info["entries"] > iu.size
    assert np.array_equal(gpu_grid(ctx, vis, data, C, gSize, "tiled", tuning), gref)
    assert np.array_equal(gpu_degrid(ctx, vis, G, C, "tiled", tuning), oref)


def test_tuning_kernel_field_validated(ctx):
    Sz, gSize = 33, 128
    iu, iv, co, data, C = small_problem(78, 50, gSize, Sz)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    with pytest.raises(tcv.TcvError):
        tcv.Plan(ctx, vis, kernel=5)
    with pytest.raises(tcv.TcvError):
        tcv.Plan(ctx, vis, degrid_kernel=7)
    info = tcv.Plan(ctx, vis, kernel=2, degrid_kernel=3, slice_keys=64, phases=16,
                    grid_phases=4).info()
    assert (info["kernel"], info["degrid_kernel"], info["phases"], info["grid_phases"]) == (2, 3, 16, 4)
    # the two-column gridder needs 4 waves x 128/tile_w rows: (8, 32) falls back
    assert tcv.Plan(ctx, vis, tile_w=8, tile_h=32, kernel=3).info()["kernel"] == 1


def test_repeat_runs_burnin_style(ctx):
    """BurnIn-style check (attic/tConvolveBurnIn/Benchmark.cc:343-374): the
    same work twice on one plan.  With every tile a single work item the tiled
    gridder has no atomics and is bit-reproducible; with split work items the
    FP32 atomic flushes may round in a different order (within 1e-6 of the
    max), and so may a rebuilt plan (its binning order is not fixed)."""
    Sz, gSize = 129, 512
    iu, iv, co, data, C = small_problem(91, 6000, gSize, Sz, cluster=150)
    vis = vis_from_host(ctx, iu, iv, co, [Sz], gSize, C.size)
    d_data, d_C = ctx.to_device(data), ctx.to_device(C)

    def run(plan):
        g = ctx.zeros((gSize, gSize), np.complex64)
        plan.grid(d_data, d_C, g)
        return g.download()

    for kernel in (1, 2, 3):
        det = tcv.Plan(ctx, vis, chunk=1 << 20, slice_keys=1, phases=1, kernel=kernel)
        a, b = run(det), run(det)
        assert np.array_equal(a, b), kernel
        split = tcv.Plan(ctx, vis, chunk=64, slice_keys=64, phases=8, kernel=kernel,
                         degrid_kernel=kernel)
        x, y = run(split), run(split)
        print(f"kernel {kernel}: {int(np.count_nonzero(x != y))} pixels differ between repeats")
        assert rel_err(x, y) <= 1e-6
        assert rel_err(x, a) <= TOL
