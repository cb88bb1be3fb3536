"""FP16 storage mode (C(w) and visibilities as half pairs, per-w-plane
max-abs scaling, FP32 accumulation): errors against the FP64 oracle and the
FP32 path, on the config-1 workload and a complex-C variable-width job."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tcv
from helpers import rel_err, small_problem
from tcv import setup as S

pytestmark = pytest.mark.gpu
# FP16 keeps 11 significant bits: each stored value is within 2^-12 of itself
# (relative, normal range).  Measured against FP64 on these seeded problems:
# 3.7e-4 (grid) and 2.0e-4 (degrid) of the maximum on the small complex job,
# 1.8e-4 / 3.9e-5 on the config-1 workload; the bar sits at 1.6x the largest.
HALF_TOL = 6e-4


def _half_run(ctx, vis, data, C, nPlanes, planeTaps, gSize, planeStart=None, kernel=0):
    """kernel: the same code for the gridder and the degridder (0 = auto)."""
    d_C, d_data = ctx.to_device(C), ctx.to_device(data)
    Ch = tcv.HalfTable(ctx, d_C, nPlanes, planeTaps, planeStart)
    dh = tcv.HalfVis(ctx, d_data)
    plan = tcv.Plan(ctx, vis, kernel=kernel, degrid_kernel=kernel)
    g = ctx.zeros((gSize, gSize), np.complex64)
    plan.grid_half(dh, Ch, g)
    out = ctx.empty(vis.nVis, np.complex64)
    plan.degrid_half(g, Ch, out)
    return g.download(), out.download(), Ch


@pytest.mark.parametrize("kernel", [1, 2, 3, 4])
def test_half_small_complex_vs_fp64(ctx, kernel):
    Sz, gSize, os_, planes = 65, 320, 8, 5
    iu, iv, co, data, C = small_problem(61, 4000, gSize, Sz, os_=os_, nPlanes=planes)
    vis = tcv.Visibilities(ctx, iu, iv, co, [Sz], gSize, C.size)
    g, out, Ch = _half_run(ctx, vis, data, C, planes, Sz * Sz * os_ * os_, gSize, kernel=kernel)
    g64 = O.grid_f64(data, iu, iv, co, C, gSize, Sz)
    e = rel_err(g, g64)
    print(f"fp16 grid err vs fp64 {e:.2e}")
    assert e <= HALF_TOL
    o64 = O.degrid_f64(g.astype(np.complex128), iu, iv, co, C, Sz)
    eo = rel_err(out, o64)
    print(f"fp16 degrid err vs fp64 {eo:.2e}")
    assert eo <= HALF_TOL
    # the scales are the per-plane max |component|
    Cp = C.reshape(planes, -1)
    assert np.allclose(Ch.host_scales, np.maximum(np.abs(Cp.real), np.abs(Cp.imag)).max(1))


def test_half_variable_width_planes(ctx):
    """Current-generation layout (per-plane widths, complex C, planes from
    planeStart and the plan's w-plane)."""
    b = tcv.CurrentBenchmark(ctx)
    b.setRunType(1)
    b.init()
    iu, iv = b.samples.iu.download(), b.samples.iv.download()
    wp, co = b.samples.wPlane.download(), b.samples.cOffset.download()
    C = b.C.download()
    n = 200000
    sel = np.random.default_rng(3).choice(iu.size, n, replace=False)
    data = ((np.random.default_rng(4).standard_normal(n) + 1j) * 0.5).astype(np.complex64)
    vis = tcv.Visibilities(ctx, iu[sel], iv[sel], co[sel], b.sSize, b.gSize, C.size, wPlane=wp[sel])
    ps = np.append(b.cOffset0.astype(np.int64), C.size)
    g, out, _ = _half_run(ctx, vis, data, C, b.wSize, 0, b.gSize, planeStart=ps)
    g64 = O.grid_f64(data, iu[sel], iv[sel], co[sel], C, b.gSize, b.sSize, wPlane=wp[sel])
    o64 = O.degrid_f64(g.astype(np.complex128), iu[sel], iv[sel], co[sel], C, b.sSize, wPlane=wp[sel])
    e, eo = rel_err(g, g64), rel_err(out, o64)
    print(f"fp16 variable widths vs fp64: grid {e:.2e}, degrid {eo:.2e}")
    assert e <= HALF_TOL and eo <= HALF_TOL


@pytest.mark.parametrize("variant", ["reference", "wide"])
def test_half_attic_errors_reported(ctx, variant):
    """The config-1 workload (160 000 samples, 4096^2 grid, width 129; the
    reference generator's [-1000, 1000] m and BASELINE's [-2000, 2000] m) in
    FP16 storage: max and RMS error of the FP32 and FP16 gridding and
    degridding against the FP64 oracle (degridding of the FP32 grid, the
    same input for all three)."""
    prob = S.attic_problem(variant)
    job = tcv.GridJob(ctx, prob)
    iu, iv, _, co = job.offsets_host()
    planeTaps = 129 * 129 * 64
    job.zero_grid()
    job.grid("tiled")
    g32 = job.grid_d.download()
    g64 = O.grid_f64(prob.data, iu, iv, co, prob.C, 4096, 129)
    Ch = tcv.HalfTable(ctx, job.C, 33, planeTaps)
    dh = tcv.HalfVis(ctx, job.data)
    gh = ctx.zeros((4096, 4096), np.complex64)
    job.plan.grid_half(dh, Ch, gh)
    g16 = gh.download()

    def errs(a, ref):
        m = np.abs(ref).max()
        return np.abs(a - ref).max() / m, np.sqrt(np.mean(np.abs(a - ref) ** 2)) / m

    e32, r32 = errs(g32, g64)
    e16, r16 = errs(g16, g64)
    job.degrid("tiled")                 # FP32 degridding of the FP32 grid
    o32 = job.out.download()
    oh = ctx.empty(prob.nVis, np.complex64)
    job.plan.degrid_half(job.grid_d, Ch, oh)
    o16 = oh.download()
    o64 = O.degrid_f64(g32.astype(np.complex128), iu, iv, co, prob.C, 129)
    d32, q32 = errs(o32, o64)
    d16, q16 = errs(o16, o64)
    print(f"config 1 ({variant}) vs FP64: grid fp32 max {e32:.2e} rms {r32:.2e}, "
          f"fp16 max {e16:.2e} rms {r16:.2e}; degrid fp32 max {d32:.2e} rms {q32:.2e}, "
          f"fp16 max {d16:.2e} rms {q16:.2e}")
    assert e32 <= 1e-5 and d32 <= 1e-5
    assert e16 <= 3e-4 and d16 <= 1e-4
    # RMS well below the maximum error: rounding, not a systematic offset
    assert r16 <= e16 / 2 and q16 <= d16 / 2


def test_half_plan_defaults_for_four_byte_taps(ctx):
    """A plan made for FP16 storage (tap_bytes=4) keeps 64-wide tiles at width
    129 and cuts the table into half as many phases as an FP32 plan (the same
    bytes per phase); its FP16 results stay within the FP16 tolerance, and a
    tap size other than 0, 4 or 8 is refused."""
    Sz, gSize, os_, planes = 129, 400, 8, 5
    iu, iv, co, data, C = small_problem(730, 3000, gSize, Sz, os_=os_, nPlanes=planes)
    vis = tcv.Visibilities(ctx, iu, iv, co, [Sz], gSize, C.size)
    i32 = tcv.Plan(ctx, vis).info()
    plan = tcv.Plan(ctx, vis, tap_bytes=4)
    i16 = plan.info()
    assert (i32["tile_w"], i32["tap_bytes"]) == (128, 8)
    assert (i16["tile_w"], i16["tile_h"], i16["tap_bytes"]) == (64, 64, 4)
    assert i16["phases"] <= i32["phases"]
    d_C, d_data = ctx.to_device(C), ctx.to_device(data)
    Ch = tcv.HalfTable(ctx, d_C, planes, Sz * Sz * os_ * os_)
    dh = tcv.HalfVis(ctx, d_data)
    g = ctx.zeros((gSize, gSize), np.complex64)
    plan.grid_half(dh, Ch, g)
    g64 = O.grid_f64(data, iu, iv, co, C, gSize, Sz)
    assert rel_err(g.download(), g64) <= HALF_TOL
    out = ctx.empty(vis.nVis, np.complex64)
    plan.degrid_half(ctx.to_device(g64.astype(np.complex64)), Ch, out)
    assert rel_err(out.download(), O.degrid_f64(g64, iu, iv, co, C, Sz)) <= HALF_TOL
    with pytest.raises(tcv.TcvError):
        tcv.Plan(ctx, vis, tap_bytes=2)


@pytest.mark.parametrize("Sz", [5, 129])
def test_half_wide_tiles(ctx, Sz):
    """FP16 storage on 128 x 64 tiles (two-tap pair kernels): within the FP16
    tolerance of the FP64 oracle; the four-tap code is refused on such tiles."""
    gSize, os_, planes = 300, 4, 3
    iu, iv, co, data, C = small_problem(720 + Sz, 1500, gSize, Sz, os_=os_, nPlanes=planes,
                                        cluster=60)
    vis = tcv.Visibilities(ctx, iu, iv, co, [Sz], gSize, C.size)
    d_C, d_data = ctx.to_device(C), ctx.to_device(data)
    Ch = tcv.HalfTable(ctx, d_C, planes, Sz * Sz * os_ * os_)
    dh = tcv.HalfVis(ctx, d_data)
    plan = tcv.Plan(ctx, vis, tile_w=128, tile_h=64, kernel=3, degrid_kernel=3, phases=2,
                    slice_keys=8, chunk=64)
    g = ctx.zeros((gSize, gSize), np.complex64)
    plan.grid_half(dh, Ch, g)
    g64 = O.grid_f64(data, iu, iv, co, C, gSize, Sz)
    assert rel_err(g.download(), g64) <= HALF_TOL
    out = ctx.empty(vis.nVis, np.complex64)
    plan.degrid_half(ctx.to_device(g64.astype(np.complex64)), Ch, out)
    assert rel_err(out.download(), O.degrid_f64(g64, iu, iv, co, C, Sz)) <= HALF_TOL
    with pytest.raises(tcv.TcvError):
        tcv.Plan(ctx, vis, tile_w=128, tile_h=64, kernel=4, degrid_kernel=4)


@pytest.mark.parametrize("Sz", [3, 5, 33])
@pytest.mark.parametrize("tw,th", [(64, 64), (32, 32), (16, 32), (8, 64)])
def test_half_four_tap_kernels_shapes(ctx, Sz, tw, th):
    """FP16 storage with the four-tap kernels (code 4) at every tile shape
    (shapes too narrow for four columns per lane, and kernels narrower than
    four taps, fall back to the two-tap kernels): within the FP16 tolerance."""
    gSize, os_, planes = 192, 4, 3
    iu, iv, co, data, C = small_problem(700 + Sz + tw, 1500, gSize, Sz, os_=os_, nPlanes=planes,
                                        cluster=50)
    vis = tcv.Visibilities(ctx, iu, iv, co, [Sz], gSize, C.size)
    d_C, d_data = ctx.to_device(C), ctx.to_device(data)
    Ch = tcv.HalfTable(ctx, d_C, planes, Sz * Sz * os_ * os_)
    dh = tcv.HalfVis(ctx, d_data)
    plan = tcv.Plan(ctx, vis, tile_w=tw, tile_h=th, kernel=4, degrid_kernel=4, phases=2,
                    slice_keys=8, chunk=64)
    g = ctx.zeros((gSize, gSize), np.complex64)
    plan.grid_half(dh, Ch, g)
    g64 = O.grid_f64(data, iu, iv, co, C, gSize, Sz)
    assert rel_err(g.download(), g64) <= HALF_TOL
    out = ctx.empty(vis.nVis, np.complex64)
    plan.degrid_half(ctx.to_device(g64.astype(np.complex64)), Ch, out)
    o64 = O.degrid_f64(g64, iu, iv, co, C, Sz)
    assert rel_err(out.download(), o64) <= HALF_TOL
