"""Serving: continuous batching with HBM-budget admission.  Each request's
result is the same alone and in a mixed stream; admission waits for memory;
requests that can never fit are refused; queued requests can be cancelled."""
from __future__ import annotations

import time

import numpy as np
import pytest

import tcv
from helpers import rel_err
from tcv import setup as S
from tcv.gridder import BatchJobs

pytestmark = pytest.mark.gpu
TOL = 1e-5
GB = 1 << 30


@pytest.fixture(scope="module")
def templates(ctx):
    bj = BatchJobs(ctx, S.uneven_batch(nJobs=6, seed=21, vis_min=3e3, vis_max=3e4, grids=(512, 1024, 2048)))
    bj.zero_grids()
    bj.run_batched(graph=False)     # resident grids for the degrid requests
    ctx.sync()
    return bj


def result_of(srv, rid, job, op):
    res = srv.wait(rid)
    assert res["status"] == 0
    sp = job["spec"]
    shape = (sp["gSize"], sp["gSize"]) if op == 0 else (sp["nVis"],)
    return res, srv.download(res, shape)


def submit(srv, job, op):
    if op == 0:
        return srv.submit(0, job["vis"], job["C"], data=job["data"])
    return srv.submit(1, job["vis"], job["C"], grid=job["grid"])


def test_results_same_alone_and_in_the_mix(ctx, templates):
    srv = tcv.Server(ctx, 8 * GB, max_inflight=3)
    jobs = templates.jobs
    alone = {}
    for j in range(len(jobs)):
        for op in (0, 1):
            rid = submit(srv, jobs[j], op)
            alone[(j, op)] = result_of(srv, rid, jobs[j], op)[1]
            srv.release(rid)
    # the tiled path agrees with the alone results
    for j in range(len(jobs)):
        g = ctx.zeros(alone[(j, 0)].shape, np.complex64)
        jobs[j]["plan"].grid(jobs[j]["data"], jobs[j]["C"], g)
        assert rel_err(alone[(j, 0)], g.download()) <= TOL
    # a Poisson stream in which every (job, op) appears several times, overlapping
    rng = np.random.default_rng(5)
    order = [(j, op) for j in range(len(jobs)) for op in (0, 1)] * 3
    rng.shuffle(order)
    gaps = rng.exponential(0.002, len(order))
    ids = []
    t0 = time.perf_counter()
    for (j, op), gap in zip(order, gaps):
        while time.perf_counter() - t0 < gap:
            pass
        t0 = time.perf_counter()
        ids.append((submit(srv, jobs[j], op), j, op))
    overlapped = 0
    for rid, j, op in ids:
        res, out = result_of(srv, rid, jobs[j], op)
        overlapped += res["launched"] > res["submitted"]
        assert rel_err(out, alone[(j, op)]) <= TOL, (j, op)
        srv.release(rid)
    srv.close()


def test_budget_refuses_requests_that_never_fit(ctx, templates):
    job = templates.jobs[0]
    srv = tcv.Server(ctx, 1 << 20)     # 1 MiB: smaller than any plan
    with pytest.raises(tcv.TcvError) as e:
        submit(srv, job, 0)
    assert e.value.code == "TCV_ERR_BUDGET"
    srv.close()


def test_admission_waits_for_memory_and_cancel(ctx, templates):
    big = max(templates.jobs, key=lambda j: j["spec"]["gSize"])
    # an arena that holds one such request at a time: the next ones wait
    probe = tcv.Server(ctx, 8 * GB, max_inflight=1)
    used = probe.footprint(0, big["vis"], big["C"], data=big["data"])
    probe.close()
    assert used > big["spec"]["gSize"] ** 2 * 8
    srv = tcv.Server(ctx, int(used * 1.5), max_inflight=2)
    ids = [submit(srv, big, 0) for _ in range(4)]
    victim = ids[-1]
    try:
        srv.cancel(victim)
        cancelled = True
    except tcv.TcvError:
        cancelled = False      # already admitted: it simply runs
    # results stay in the arena until released, so release each one as it
    # completes: that is what lets the next request in
    res = []
    for r in ids:
        res.append(srv.wait(r, timeout=60))
        srv.release(r)
    done = [r for r in res if r["status"] == 0]
    for a, b in zip(done, done[1:]):
        # one at a time: a request is admitted only after the previous finished
        assert b["admitted"] >= a["completed"] - 1e-4
    assert len(done) >= 3
    if cancelled:
        assert res[-1]["status"] == 7     # TCV_ERR_CANCELLED
    srv.close()


def _batches_one_slot(res):
    """With one slot, batches run one after another: a request opens a new
    batch when it launched after the running batch completed."""
    batches, done = [], -1.0
    for r in res:
        if r["launched"] >= done:
            batches.append([])
            done = r["completed"]
        batches[-1].append(r)
    return batches


@pytest.mark.parametrize("limit_of", ["below_every_request", "twice_the_median"])
def test_batch_points_limit(ctx, templates, limit_of):
    """maxBatchPoints: a batch takes admitted requests in order until the next
    would carry it past the limit (a lone request always runs); results are
    unchanged."""
    jobs = templates.jobs
    pts = sorted(j["spec"]["nVis"] * j["spec"]["sSize"] ** 2 for j in jobs)
    limit = 1 if limit_of == "below_every_request" else 2 * pts[len(pts) // 2]
    srv = tcv.Server(ctx, 8 * GB, max_inflight=1, max_batch_points=limit)
    order = [(j, op) for op in (0, 1) for j in range(len(jobs))] * 2
    ids = [(submit(srv, jobs[j], op), j, op) for j, op in order]
    res = []
    for rid, j, op in ids:
        r, out = result_of(srv, rid, jobs[j], op)
        res.append(r)
        sp = jobs[j]["spec"]
        if op == 0:
            want = ctx.zeros((sp["gSize"], sp["gSize"]), np.complex64)
            jobs[j]["plan"].grid(jobs[j]["data"], jobs[j]["C"], want)
        else:
            want = ctx.empty(sp["nVis"], np.complex64)
            jobs[j]["plan"].degrid(jobs[j]["grid"], jobs[j]["C"], want)
        assert rel_err(out, want.download()) <= TOL, (j, op)
        srv.release(rid)
    batches = _batches_one_slot(res)
    for b in batches:
        assert len(b) == 1 or sum(r["points"] for r in b) <= limit
    if limit == 1:
        assert len(batches) == len(order)
    srv.close()


def test_large_requests_run_alone(ctx, templates):
    """exclusivePoints: a request of at least that many grid points waits for
    the running batches to finish and runs with no other batch beside it,
    while smaller ones still share the device over four slots.  Every result
    equals its job run alone; the launch and completion times show each large
    request's run overlapping no other request's."""
    jobs = templates.jobs
    pts = sorted(j["spec"]["nVis"] * j["spec"]["sSize"] ** 2 for j in jobs)
    big = pts[len(pts) // 2]                   # the larger half runs alone
    srv = tcv.Server(ctx, 8 * GB, max_inflight=4, max_batch_points=2 * pts[-1], exclusive_points=big)
    order = [(j, op) for _ in range(3) for j in range(len(jobs)) for op in (0, 1)]
    ids = [(submit(srv, jobs[j], op), j, op) for j, op in order]
    res = []
    for rid, j, op in ids:
        r, out = result_of(srv, rid, jobs[j], op)
        res.append(r)
        sp = jobs[j]["spec"]
        if op == 0:
            want = ctx.zeros((sp["gSize"], sp["gSize"]), np.complex64)
            jobs[j]["plan"].grid(jobs[j]["data"], jobs[j]["C"], want)
        else:
            want = ctx.empty(sp["nVis"], np.complex64)
            jobs[j]["plan"].degrid(jobs[j]["grid"], jobs[j]["C"], want)
        assert rel_err(out, want.download()) <= TOL, (j, op)
        srv.release(rid)
    large = [r for r in res if r["points"] >= big]
    assert large
    for a in large:
        for b in res:
            if b is a:
                continue
            assert b["completed"] <= a["launched"] or b["launched"] >= a["completed"], (a, b)
    srv.close()


def test_server_config_rejects_negative_limits(ctx):
    with pytest.raises(tcv.TcvError):
        tcv.Server(ctx, GB, exclusive_points=-1)
    with pytest.raises(tcv.TcvError):
        tcv.Server(ctx, GB, max_batch_points=-1)
    with pytest.raises(tcv.TcvError):
        tcv.Server(ctx, GB, max_batch=-2)
