"""Autotuner: picks among candidate tunings by timing them on the job, caches
the winner for the workload class, and plans made afterwards without an
explicit tuning use it; results stay equal to the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import tcv
from helpers import rel_err, small_problem

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture()
def own_ctx():
    # a context of its own: the tuning cache lives in the context
    c = tcv.Context(0)
    yield c
    c.close()


def _job(c, seed=7, Sz=65, gSize=512, nVis=4000, cluster=120):
    iu, iv, co, data, C = small_problem(seed, nVis, gSize, Sz, cluster=cluster)
    vis = tcv.Visibilities(c, iu, iv, co, [Sz], gSize, C.size)
    return (iu, iv, co, data, C, Sz, gSize), vis


def test_autotune_candidates_and_cache(own_ctx):
    c = own_ctx
    (iu, iv, co, data, C, Sz, gSize), vis = _job(c)
    d_data, d_C = c.to_device(data), c.to_device(C)
    scratch_grid = c.zeros((gSize, gSize), np.complex64)
    scratch_out = c.empty(vis.nVis, np.complex64)
    cands = [dict(tile_w=64, tile_h=64, chunk=256, slice_keys=64, phases=4, xcd_split=0),
             dict(tile_w=32, tile_h=32, chunk=128, slice_keys=1, phases=1, xcd_split=0),
             dict(tile_w=64, tile_h=32, chunk=512, slice_keys=256, phases=16, xcd_split=1)]
    best = tcv.autotune(c, vis, d_data, d_C, scratch_grid, scratch_out, op=3, candidates=cands)
    assert best["ms"] > 0
    key = {k: best[k] for k in ("tile_w", "tile_h", "chunk", "slice_keys", "phases", "xcd_split")}
    assert key in cands
    # a plan made without a tuning takes the cached winner
    plan = tcv.Plan(c, vis)
    info = plan.info()
    for k in ("tile_w", "tile_h", "chunk", "phases"):
        assert info[k] == best[k], k
    g = c.zeros((gSize, gSize), np.complex64)
    plan.grid(d_data, d_C, g)
    gref = O.grid(data, iu, iv, co, C, gSize, Sz)
    assert rel_err(g.download(), gref) <= TOL
    out = c.empty(vis.nVis, np.complex64)
    plan.degrid(c.to_device(gref), d_C, out)
    assert rel_err(out.download(), O.degrid(gref, iu, iv, co, C, Sz)) <= TOL


def test_autotune_builtin_set_grid_only(own_ctx):
    c = own_ctx
    (iu, iv, co, data, C, Sz, gSize), vis = _job(c, seed=8, Sz=33, gSize=256, nVis=3000, cluster=60)
    d_data, d_C = c.to_device(data), c.to_device(C)
    g = c.zeros((gSize, gSize), np.complex64)
    best = tcv.autotune(c, vis, d_data, d_C, g, None, op=1)
    assert best["tile_w"] in (8, 16, 32, 64) and best["tile_h"] in (32, 64)
    plan = tcv.Plan(c, vis)
    g2 = c.zeros((gSize, gSize), np.complex64)
    plan.grid(d_data, d_C, g2)
    assert rel_err(g2.download(), O.grid(data, iu, iv, co, C, gSize, Sz)) <= TOL
