"""Full-size correctness on exactly what bench.py times:

* the spectral workload (1 048 576 samples x 8 channels, 4096^2 grid, width
  129, wSize 129: 1.4e11 grid points a pass) on the default plan (128 x 64
  tiles, 288 C-slice phases, 2048 keys, per-XCD queues);
* the uneven 64-job batch, binned, gridded and degridded in one captured
  graph as the bench runs it;
* served requests over the same 64 jobs at the serving workload's scale.

Bars (max |a - ref| / max |ref|, real and imaginary parts separately):
* every cell of every FP32 grid against the device's FP64 reference mode of
  the plain kernels (a different code path: one workgroup per visibility,
  FP64 atomics) at 1e-5, and every degridded visibility likewise;
* windows of 128 x 64 cells (tile-aligned, straddling tile corners, on the
  populated region's corners, on the empty grid edge) against the oracle's
  windowed gridding, which equals the oracle's full grid cell for cell: FP32
  at 1e-5, and the device FP64 grid against the oracle's FP64 window at 1e-12;
* sampled visibilities of every degridding against the oracle;
* the FP64 grid conserves the total, sum(grid) = sum_v d_v sum(slice_v), and
  the FP64 kernels are adjoint, <G, grid(d)> = <degrid(G), d>, each to 1e-9 of
  the magnitude scale (a missing visibility moves either by ~1e-3 of it)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tcv
from helpers import rel_err
from tcv import setup as S
from tcv.gridder import BatchJobs

pytestmark = pytest.mark.gpu
TOL = 1e-5
TOL64 = 1e-12


def fp64_grid(ctx, vis, data, C, gSize):
    g = ctx.zeros((gSize, gSize), np.complex128)
    tcv.grid_plain_f64(ctx, vis, data, C, g)
    out = g.download()
    g.free()
    return out


def fp64_degrid(ctx, vis, grid, C):
    out = ctx.empty(vis.nVis, np.complex128)
    tcv.degrid_plain_f64(ctx, vis, grid, C, out)
    h = out.download()
    out.free()
    return h


@pytest.fixture(scope="module")
def spectral(ctx):
    prob = S.spectral_problem()
    job = tcv.GridJob(ctx, prob)          # the default plan, as bench.py builds it
    assert job.nBad == 0
    iu, iv, _, co = job.offsets_host()
    job.zero_grid()
    job.grid("tiled")
    g32 = job.grid_d.download()
    g64 = fp64_grid(ctx, job.vis, job.data, job.C, prob.gSize)
    yield dict(prob=prob, job=job, iu=iu, iv=iv, co=co, g32=g32, g64=g64)
    job.plan.close()


def test_fullsize_default_plan(spectral):
    prob, job = spectral["prob"], spectral["job"]
    info = job.plan.info()
    assert (info["tile_w"], info["tile_h"], info["kernel"], info["degrid_kernel"]) == (128, 64, 3, 3)
    assert info["out_of_range"] == 0 and info["entries"] > prob.nVis
    assert info["xcd_split"] == 1 and info["phases"] >= 8


def test_fullsize_grid_every_cell_vs_fp64(spectral):
    e = rel_err(spectral["g32"], spectral["g64"])
    print(f"spectral grid: max error vs FP64 {e:.2e} of the maximum")
    assert e <= TOL


# (x0, y0) of 128 x 64 windows.  The populated region is the square of
# half-width f7 * 6000 / 15 = 1877 pixels about the centre; footprints reach
# 64 columns left of and 129 rows below their visibility's pixel.
WINDOWS = [
    (2048, 2048, "tile-aligned at the centre"),
    (2048 - 64, 2048 - 32, "straddling a tile corner at the centre"),
    (1037, 1505, "straddling tile edges off-centre"),
    (100, 160, "the populated square's lower-left corner"),
    (3900, 3990, "the populated square's upper-right corner"),
    (0, 2000, "the grid's left edge, partly populated"),
    (4096 - 128, 0, "the grid's corner, empty"),
]


@pytest.mark.parametrize("x0,y0,what", WINDOWS, ids=[w[2] for w in WINDOWS])
def test_fullsize_grid_windows_vs_oracle(spectral, x0, y0, what):
    p = spectral["prob"]
    w, h = 128, 64
    args = (p.data, spectral["iu"], spectral["iv"], spectral["co"], p.C, p.gSize, int(p.sSize[0]),
            x0, y0, w, h)
    o32 = O.grid_window(*args)
    o64 = O.grid_window(*args, f64=True)
    got32 = spectral["g32"][y0:y0 + h, x0:x0 + w]
    got64 = spectral["g64"][y0:y0 + h, x0:x0 + w]
    if not o64.any():
        assert not got32.any() and not got64.any(), what
        return
    assert rel_err(got32, o32) <= TOL, what
    assert rel_err(got32, o64) <= TOL, what
    assert rel_err(got64, o64) <= TOL64, what


def test_fullsize_fp64_grid_conserves_total(spectral):
    p = spectral["prob"]
    s2 = int(p.sSize[0]) ** 2
    C = p.C.reshape(-1, s2)
    idx = spectral["co"] // s2
    assert (spectral["co"] % s2 == 0).all()
    ssum = C.sum(axis=1, dtype=np.complex128)[idx]
    sabs = np.abs(C).sum(axis=1, dtype=np.float64)[idx]
    d = p.data.astype(np.complex128)
    want = (d * ssum).sum()
    scale = (np.abs(d) * sabs).sum()
    total = spectral["g64"].sum()
    # one visibility contributes |d| |sum(slice)| ~ 0.1 .. 1 to the total
    assert abs(total - want) <= 1e-9 * scale, (total, want, scale)


def test_fullsize_degrid_every_visibility_vs_fp64_and_oracle(ctx, spectral):
    p, job = spectral["prob"], spectral["job"]
    job.plan.degrid(job.grid_d, job.C, job.out)      # degrid the gridded grid
    out32 = job.out.download()
    out64 = fp64_degrid(ctx, job.vis, job.grid_d, job.C)
    e = rel_err(out32, out64)
    print(f"spectral degrid: max error vs FP64 {e:.2e} of the maximum")
    assert e <= TOL
    rng = np.random.default_rng(43)
    pick = np.sort(rng.choice(p.nVis, 4000, replace=False))
    iu, iv, co = spectral["iu"][pick], spectral["iv"][pick], spectral["co"][pick]
    s = int(p.sSize[0])
    g32 = spectral["g32"]
    assert rel_err(out32[pick], O.degrid(g32, iu, iv, co, p.C, s)) <= TOL
    assert rel_err(out64[pick], O.degrid_f64(g32, iu, iv, co, p.C, s)) <= TOL64


def test_fullsize_degrid_of_ones_is_slice_sum(ctx, spectral):
    p, job = spectral["prob"], spectral["job"]
    s2 = int(p.sSize[0]) ** 2
    C = p.C.reshape(-1, s2)
    idx = spectral["co"] // s2
    ssum = C.sum(axis=1, dtype=np.complex128)[idx]
    sabs = np.abs(C).sum(axis=1, dtype=np.float64)[idx]
    ones = ctx.to_device(np.ones((p.gSize, p.gSize), np.complex64))
    job.plan.degrid(ones, job.C, job.out)
    out = job.out.download().astype(np.complex128)
    err = np.abs(out - ssum) / np.maximum(sabs, 1e-30)
    assert err.max() <= TOL, float(err.max())


def test_fullsize_fp64_kernels_adjoint(ctx, spectral):
    p, job = spectral["prob"], spectral["job"]
    rng = np.random.default_rng(41)
    g = p.gSize
    G = (rng.standard_normal((g, g), dtype=np.float32) +
         1j * rng.standard_normal((g, g), dtype=np.float32)).astype(np.complex64)
    dG = ctx.to_device(G)
    dg = fp64_degrid(ctx, job.vis, dG, job.C)
    d = p.data.astype(np.complex128)
    lhs = (G.astype(np.complex128) * spectral["g64"]).sum()
    rhs = (dg * d).sum()
    s2 = int(p.sSize[0]) ** 2
    sabs = np.abs(p.C.reshape(-1, s2)).sum(axis=1, dtype=np.float64)[spectral["co"] // s2]
    scale = (np.abs(d) * sabs).sum() * float(np.abs(G).max())
    assert abs(lhs - rhs) <= 1e-9 * scale, (lhs, rhs, scale)
    # and the FP32 tiled degridder of the same random grid, every visibility
    job.plan.degrid(dG, job.C, job.out)
    assert rel_err(job.out.download(), dg) <= TOL


# ---- the 64-job uneven batch and serving at their benchmark sizes ----------------

@pytest.fixture(scope="module")
def batch(ctx):
    specs = S.uneven_batch(nJobs=64)      # the bench's mix (seed 3)
    bj = BatchJobs(ctx, specs)             # the bench's tuning (shared tile rule)
    bj.zero_grids()
    bj.run_batched(graph=True)             # re-bin + grid + degrid in one captured graph
    ctx.sync()
    yield bj


def test_batch_as_benched_every_job(ctx, batch):
    """Every job of the batch: every grid cell and every degridded visibility
    against the FP64 reference kernels, a 128 x 64 window straddling the
    centre tiles against the oracle, and 300 sampled visibilities of the
    degridding against the oracle."""
    rng = np.random.default_rng(44)
    worst = [0.0, 0.0]
    for j, job in enumerate(batch.jobs):
        sp, vis = job["spec"], job["vis"]
        g, s = sp["gSize"], sp["sSize"]
        g32 = job["grid"].download()
        e = rel_err(g32, fp64_grid(ctx, vis, job["data"], job["C"], g))
        assert e <= TOL, (j, e)
        worst[0] = max(worst[0], e)
        out64 = fp64_degrid(ctx, vis, job["grid"], job["C"])
        out32 = job["out"].download()
        e = rel_err(out32, out64)
        assert e <= TOL, (j, e)
        worst[1] = max(worst[1], e)
        iu, iv, co = vis.iu.download(), vis.iv.download(), vis.cOffset.download()
        C = job["C"].download()
        x0, y0 = g // 2 - 64, g // 2 - 32
        win = O.grid_window(sp["data"], iu, iv, co, C, g, s, x0, y0, 128, 64)
        assert rel_err(g32[y0:y0 + 64, x0:x0 + 128], win) <= TOL, j
        pick = np.sort(rng.choice(sp["nVis"], min(300, sp["nVis"]), replace=False))
        want = O.degrid(g32, iu[pick], iv[pick], co[pick], C, s)
        assert rel_err(out32[pick], want) <= TOL, j
        del C
        if j % 8 == 7:
            print(f"batch job {j}: checked", flush=True)
    print(f"batch: max error vs FP64 grid {worst[0]:.2e}, degrid {worst[1]:.2e}")


def test_served_requests_at_scale(ctx, batch):
    """A Poisson stream of 400 requests over the 64 resident jobs (50 % grid
    into new grids, 50 % degrid of the resident grids) through a server whose
    arena is what the 200 GB budget leaves: every result equals the batched
    result of its job within 1e-5, and 16 of them also the oracle (a window of
    each grid, sampled values of each degridding)."""
    import time
    from tcv.gridder import batch_tile_w
    specs = [j["spec"] for j in batch.jobs]
    free, total = ctx.mem_info()
    arena = int(min(200 * 10**9 - (total - free), free - 16 * 2**30))
    srv = tcv.Server(ctx, arena, max_inflight=4, max_batch_points=int(1e10),
                     tile_w=batch_tile_w([sp["sSize"] for sp in specs]), tile_h=64)
    rng = np.random.default_rng(45)
    n = 400
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, n))
    ops = rng.integers(0, 2, n)
    which = rng.integers(0, len(batch.jobs), n)
    ids = []
    t0 = time.perf_counter()
    for k in range(n):
        while time.perf_counter() - t0 < arrivals[k]:
            time.sleep(0.0005)
        job = batch.jobs[which[k]]
        if ops[k] == 0:
            ids.append(srv.submit(0, job["vis"], job["C"], data=job["data"]))
        else:
            ids.append(srv.submit(1, job["vis"], job["C"], grid=job["grid"]))
    checked = 0
    for k, rid in enumerate(ids):
        res = srv.wait(rid, timeout=600)
        assert res["status"] == 0, k
        job = batch.jobs[which[k]]
        sp = job["spec"]
        if ops[k] == 0:
            got = srv.download(res, (sp["gSize"], sp["gSize"]))
            assert rel_err(got, job["grid"].download()) <= TOL, k
        else:
            got = srv.download(res, (sp["nVis"],))
            assert rel_err(got, job["out"].download()) <= TOL, k
        if k % 25 == 0:
            iu, iv, co = job["vis"].iu.download(), job["vis"].iv.download(), job["vis"].cOffset.download()
            C = job["C"].download()
            g, s = sp["gSize"], sp["sSize"]
            if ops[k] == 0:
                x0, y0 = g // 2 - 200, g // 2 + 37
                win = O.grid_window(sp["data"], iu, iv, co, C, g, s, x0, y0, 128, 64)
                assert rel_err(got[y0:y0 + 64, x0:x0 + 128], win) <= TOL, k
            else:
                pick = np.sort(rng.choice(sp["nVis"], min(200, sp["nVis"]), replace=False))
                want = O.degrid(job["grid"].download(), iu[pick], iv[pick], co[pick], C, s)
                assert rel_err(got[pick], want) <= TOL, k
            checked += 1
        srv.release(rid)
    srv.close()
    assert checked == 16
