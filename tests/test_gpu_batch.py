"""Batching: many uneven jobs (different grid sizes, kernel widths, w-plane
counts and visibility counts) in one persistent launch, optionally captured
in a hipGraph, checked job by job against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tcv
from helpers import exact_problem, rel_err
from tcv import setup as S
from tcv.gridder import BatchJobs

pytestmark = pytest.mark.gpu
TOL = 1e-5


def small_specs(n=5, seed=11):
    return S.uneven_batch(nJobs=n, seed=seed, vis_min=2e3, vis_max=2e4, grids=(512, 1024, 2048))


def oracle_job(bj, j, grid_in=None):
    sp = bj.jobs[j]["spec"]
    vis = bj.jobs[j]["vis"]
    iu, iv, co = vis.iu.download(), vis.iv.download(), vis.cOffset.download()
    C = bj.jobs[j]["C"].download()
    g = O.grid(sp["data"], iu, iv, co, C, sp["gSize"], sp["sSize"])
    o = O.degrid(g if grid_in is None else grid_in, iu, iv, co, C, sp["sSize"])
    return g, o


@pytest.mark.parametrize("graph", [False, True])
def test_batch_vs_oracle(ctx, graph):
    bj = BatchJobs(ctx, small_specs())
    bj.zero_grids()
    bj.run_batched(graph=graph)
    for j in range(len(bj.jobs)):
        g = bj.jobs[j]["grid"].download()
        gref, _ = oracle_job(bj, j)
        assert rel_err(g, gref) <= TOL, j
        # degrid compared on the GPU's own grid (same input both sides)
        _, oref = oracle_job(bj, j, grid_in=g)
        assert rel_err(bj.jobs[j]["out"].download(), oref) <= TOL, j


def test_batch_modes_agree(ctx):
    bj = BatchJobs(ctx, small_specs(6, seed=12))
    res = {}
    for mode in ("plain", "tiled", "batched"):
        bj.zero_grids()
        getattr(bj, "run_" + mode)()
        res[mode] = [(j["grid"].download(), j["out"].download()) for j in bj.jobs]
    for mode in ("tiled", "batched"):
        for (g, o), (gp, op) in zip(res[mode], res["plain"]):
            assert rel_err(g, gp) <= TOL
            assert rel_err(o, op) <= TOL


def test_batch_graph_replays(ctx):
    """Replays of the captured graph, with eager work and synchronisations in
    between, each equal to the eager batched result."""
    bj = BatchJobs(ctx, small_specs(4, seed=13))
    bj.zero_grids()
    bj.run_batched(graph=False)
    eager = [(j["grid"].download(), j["out"].download()) for j in bj.jobs]
    for rep in range(3):
        bj.zero_grids()
        bj.run_batched(graph=True)
        ctx.sync()
        got = [(j["grid"].download(), j["out"].download()) for j in bj.jobs]
        for (g, o), (ge, oe) in zip(got, eager):
            assert rel_err(g, ge) <= TOL, rep
            assert rel_err(o, oe) <= TOL, rep
        # unrelated eager work between replays
        j0 = bj.jobs[0]
        j0["plan"].grid(j0["data"], j0["C"], j0["grid"])
        ctx.sync()


def test_device_initC_matches_host_taps(ctx):
    """Device-generated attic taps equal the host table up to the
    normalisation factor (FP64 sum on the device vs the reference's float
    running sum) and last-ulp libm differences."""
    f0 = S.F0
    sup, wcs, dC = tcv.attic_initC_device(ctx, f0, 5.0, 33, 2000)
    s2, w2, hC = S.attic_initC(f0, 5.0, 33, 2000)
    assert sup == s2 and wcs == w2
    d = dC.download()
    assert not d.imag.any()
    # compare shapes after removing each table's own normalisation
    ratio = d.real.astype(np.float64) / np.where(hC.real != 0, hC.real, 1)
    k = np.abs(hC.real) > 1e-3 * np.abs(hC.real).max()
    r = np.median(ratio[k])
    assert np.abs(ratio[k] / r - 1).max() < 1e-5
    # the device table is normalised to sum |C| = wSize * os^2 in FP64
    assert np.abs(d.astype(np.complex128)).sum() == pytest.approx(33 * 64, rel=1e-6)


@pytest.mark.parametrize("kernels", [(2, 2, 64), (3, 3, 64), (1, 3, 64), (3, 3, 128)])
def test_batch_other_kernels_vs_oracle(ctx, kernels):
    # gridder and degridder items cut at different phase counts; 128 x 64 tiles
    # (pair kernels) through the batched launches too
    bj = BatchJobs(ctx, small_specs(4, seed=21), tuning=dict(kernel=kernels[0],
                                                              degrid_kernel=kernels[1],
                                                              tile_w=kernels[2],
                                                              phases=8, grid_phases=2))
    bj.zero_grids()
    bj.run_batched(graph=True)
    for j in range(len(bj.jobs)):
        g = bj.jobs[j]["grid"].download()
        gref, _ = oracle_job(bj, j)
        assert rel_err(g, gref) <= TOL, j
        _, oref = oracle_job(bj, j, grid_in=g)
        assert rel_err(bj.jobs[j]["out"].download(), oref) <= TOL, j


@pytest.mark.parametrize("tile_w", [64, 128])
def test_batch_jobs_sharing_one_grid(ctx, tile_w):
    """Two jobs gridding into ONE grid in one persistent launch (two channels
    into one image).  Each job's plan has one C-slice phase and work items
    larger than any tile's entry list, so within its own job every item holds
    its tile alone; the batch must still flush those tiles with atomics, since
    the other job's item for the same tile runs at the same time.  Exact
    inputs: the result equals the oracle's sum of both jobs bit for bit (a
    lost read-modify-write would show as a whole missing footprint piece)."""
    Sz, gSize = 129, 512
    jobs = [exact_problem(seed, 600, gSize, Sz, cluster=None) for seed in (31, 32)]
    C = jobs[0][4]
    d_C = ctx.to_device(C)
    grid = ctx.zeros((gSize, gSize), np.complex64)
    keep = []
    ref = np.zeros((gSize, gSize), np.complex64)
    for iu, iv, co, data, _, _ in jobs:
        vis = tcv.Visibilities(ctx, iu, iv, co, [Sz], gSize, C.size)
        plan = tcv.Plan(ctx, vis, tile_w=tile_w, tile_h=64, phases=1, grid_phases=1,
                        slice_keys=1, chunk=1 << 20)
        info = plan.info()
        # one work item per non-empty tile: every item of the job is exclusive
        assert info["items"] == len({(u // tile_w, v // 64)
                                     for u0, v0 in zip(iu - Sz // 2, iv)
                                     for u in range(u0 // tile_w * tile_w, u0 + Sz, tile_w)
                                     for v in range(v0 // 64 * 64, v0 + Sz, 64)})
        keep.append((vis, plan, ctx.to_device(data)))
        ref = O.grid(data, iu, iv, co, C, gSize, Sz, out=ref)
    b = tcv.Batch(ctx, [(plan, d, d_C, grid, None) for vis, plan, d in keep])
    for rep in range(3):
        grid.zero()
        b.grid()
        ctx.sync()
        assert np.array_equal(grid.download(), ref), rep


@pytest.mark.parametrize("tw,th,ie,cols", [(64, 128, 32, 0), (32, 256, 16, 0), (128, 64, 8, 0),
                                          (128, 64, 32, 32), (128, 64, 24, 128), (64, 128, 32, 16)])
def test_batch_tall_tiles_and_fine_phases(ctx, tw, th, ie, cols):
    """Tall tiles (pair kernels only), automatic phase counts that keep as
    few as `ie` entries per work item and the pair gridder's column groups,
    through the captured batch: every job against the oracle."""
    bj = BatchJobs(ctx, small_specs(4, seed=23),
                   tuning=dict(tile_w=tw, tile_h=th, kernel=3, degrid_kernel=3, item_entries=ie,
                               grid_cols=cols))
    info = bj.jobs[0]["plan"].info()
    assert (info["tile_w"], info["tile_h"], info["item_entries"]) == (tw, th, ie)
    assert info["grid_cols"] == (cols or (64 if tw == 128 else tw))
    bj.zero_grids()
    bj.run_batched(graph=True)
    for j in range(len(bj.jobs)):
        g = bj.jobs[j]["grid"].download()
        gref, _ = oracle_job(bj, j)
        assert rel_err(g, gref) <= TOL, j
        _, oref = oracle_job(bj, j, grid_in=g)
        assert rel_err(bj.jobs[j]["out"].download(), oref) <= TOL, j
