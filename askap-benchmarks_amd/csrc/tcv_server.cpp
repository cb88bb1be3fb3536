// Serving loop: continuous batching of gridding / degridding requests on one
// device.  A scheduler thread admits queued requests first-in first-out as
// the HBM arena (the memory budget) can hold their buffers, forms a batch of
// every admitted request as soon as a stream slot is free, and launches it as
// one persistent multi-job kernel (re-binning included) on that slot's
// stream.  Several slots keep batches overlapping on the device; completion is
// detected by polling each slot's event.  Request buffers (plan, new grid or
// output) live in the arena, so the loop never calls hipMalloc or hipFree.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "arena.h"
#include "tcv_internal.h"

namespace {

enum ReqState { QUEUED = 0, ADMITTED = 1, RUNNING = 2, DONE = 3, CANCELLED = 4, FAILED = 5 };

struct Req {
    long long id = 0;
    tcv_request r{};
    std::vector<int> sSize;
    int state = QUEUED;
    tcv_status status = TCV_OK;
    double tSub = 0, tAdm = 0, tLaunch = 0, tDone = 0;
    tcv_plan plan;
    size_t planBytes = 0, planOff = 0, resBytes = 0, resOff = 0;
    bool hasPlan = false, hasRes = false;
    long long points = 0;
};

struct Slot {
    tcv_ctx *ctx = nullptr;
    tcv_batch grid, degrid;
    JobDesc *pinned = nullptr;     // host descriptors (grid jobs then degrid jobs)
    hipEvent_t done = nullptr;
    std::vector<long long> reqs;
    bool busy = false;
    bool alone = false;            // the batch holds a request of exclusivePoints or more
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct tcv_server {
    tcv_ctx *ctx = nullptr;
    tcv_server_config cfg{};
    tcv_tuning tuning{};
    char *arenaBase = nullptr;
    tcv_detail::Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv;
    std::map<long long, std::unique_ptr<Req>> reqs;
    std::deque<long long> waiting, admitted;
    long long nextId = 0;
    std::atomic<bool> stop{false};
    std::thread th;
    Clock::time_point t0;
    std::string err;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void release_plan(tcv_server *s, Req &r)
{
    if (r.hasPlan) s->arena.release(r.planOff, r.planBytes);
    r.hasPlan = false;
}

void release_result(tcv_server *s, Req &r)
{
    if (r.hasRes) s->arena.release(r.resOff, r.resBytes);
    r.hasRes = false;
}

// caller holds the lock
void finish(tcv_server *s, Req &r, int state, tcv_status st)
{
    r.state = state;
    r.status = st;
    r.tDone = s->now();
    release_plan(s, r);
    if (state != DONE || s->cfg.autoRelease) release_result(s, r);
}

// Enqueues one batch on a free slot; the lock is not held (the requests are
// ADMITTED and only this thread touches them).
tcv_status launch_slot(tcv_server *s, Slot &sl, std::vector<Req *> &batch)
{
    tcv_ctx *c = sl.ctx;
    std::vector<JobDesc> g, d;
    for (Req *r : batch) {
        tcv_plan &p = r->plan;
        TCV_HIP(c, hipMemcpyAsync(p.d_sSize, r->sSize.data(), sizeof(int) * r->sSize.size(),
                                  hipMemcpyHostToDevice, c->stream));
        const tcv_geometry &G = r->r.geom;
        GeomDev gd;
        gd.nVis = G.nVis;
        gd.iu = G.d_iu;
        gd.iv = G.d_iv;
        gd.cOffset = G.d_cOffset;
        gd.wPlane = G.d_wPlane;
        gd.nPlanes = G.nPlanes;
        gd.sUniform = r->sSize[0];
        gd.gSize = G.gSize;
        gd.cSize = G.cSize;
        gd.sSize = G.d_wPlane ? p.d_sSize : nullptr;
        tcv_status st = plan_build_kernels(c, &p, gd);
        if (st != TCV_OK) return st;
        float2 *res = (float2 *)(s->arenaBase + r->resOff);
        if (r->r.op == 0) {
            st = launch_fill_u32(c, res, 0, (size_t)G.gSize * G.gSize * 2);
            if (st != TCV_OK) return st;
            g.push_back(make_job_desc(&p, true, r->r.d_data, r->r.d_C, res, nullptr, nullptr));
        } else {
            d.push_back(make_job_desc(&p, false, nullptr, r->r.d_C, (float2 *)r->r.d_grid, res, nullptr));
        }
    }
    size_t k = 0;
    auto run = [&](tcv_batch &b, std::vector<JobDesc> &v, bool degrid) -> tcv_status {
        if (v.empty()) return TCV_OK;
        b.nJobs = (int)v.size();
        // one descriptor array per slot batch: the grid batch reads it as the
        // gridder's, the degrid batch as the degridder's
        (degrid ? b.hostD : b.host) = v;
        b.d_jobsD = b.d_jobs;
        std::memcpy(sl.pinned + k, v.data(), sizeof(JobDesc) * v.size());
        TCV_HIP(c, hipMemcpyAsync(b.d_jobs, sl.pinned + k, sizeof(JobDesc) * v.size(),
                                  hipMemcpyHostToDevice, c->stream));
        k += v.size();
        return launch_batch(c, &b, degrid);
    };
    tcv_status st = run(sl.grid, g, false);
    if (st == TCV_OK) st = run(sl.degrid, d, true);
    if (st != TCV_OK) return st;
    TCV_HIP(c, hipEventRecord(sl.done, c->stream));
    return TCV_OK;
}

void scheduler(tcv_server *s)
{
    (void)hipSetDevice(s->ctx->device);
    while (!s->stop.load()) {
        bool did = false;
        // completions
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            std::lock_guard<std::mutex> lk(s->mu);
            sl.alone = false;
            for (long long id : sl.reqs) {
                auto it = s->reqs.find(id);
                if (it != s->reqs.end())
                    finish(s, *it->second, q == hipSuccess ? DONE : FAILED,
                           q == hipSuccess ? TCV_OK : TCV_ERR_DEVICE);
            }
            if (q != hipSuccess) s->err = hipGetErrorString(q);
            sl.reqs.clear();
            sl.busy = false;
            did = true;
            s->cv.notify_all();
        }
        // admission, first in first out, as the arena frees up
        {
            std::lock_guard<std::mutex> lk(s->mu);
            while (!s->waiting.empty()) {
                auto it = s->reqs.find(s->waiting.front());
                if (it == s->reqs.end() || it->second->state != QUEUED) {
                    s->waiting.pop_front();
                    continue;
                }
                Req &r = *it->second;
                if (!s->arena.alloc(r.planBytes, &r.planOff)) break;
                if (!s->arena.alloc(r.resBytes, &r.resOff)) {
                    s->arena.release(r.planOff, r.planBytes);
                    break;
                }
                r.hasPlan = r.hasRes = true;
                plan_bind(&r.plan, s->arenaBase + r.planOff);
                r.state = ADMITTED;
                r.tAdm = s->now();
                s->admitted.push_back(r.id);
                s->waiting.pop_front();
                did = true;
            }
        }
        // launch every admitted request on free slots; a request of
        // exclusivePoints or more waits for the device to drain and runs alone
        const long long big = s->cfg.exclusivePoints;
        for (Slot &sl : s->slots) {
            if (sl.busy) continue;
            bool busyAny = false, aloneAny = false;
            for (const Slot &o : s->slots) {
                busyAny = busyAny || o.busy;
                aloneAny = aloneAny || (o.busy && o.alone);
            }
            if (aloneAny) break;
            std::vector<Req *> batch;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                long long points = 0;
                if (big > 0 && !s->admitted.empty()) {
                    Req *h = s->reqs[s->admitted.front()].get();
                    if (h->points >= big) {
                        if (busyAny) break;   // drain first
                        s->admitted.pop_front();
                        h->state = RUNNING;
                        h->tLaunch = s->now();
                        batch.push_back(h);
                        sl.reqs.push_back(h->id);
                        sl.alone = true;
                    }
                }
                while (!sl.alone && !s->admitted.empty() &&
                       (s->cfg.maxBatch <= 0 || (int)batch.size() < s->cfg.maxBatch)) {
                    Req *r = s->reqs[s->admitted.front()].get();
                    // a large request is left for a batch of its own
                    if (big > 0 && r->points >= big) break;
                    // a batch completes as one: keep large requests apart from small ones
                    if (!batch.empty() && s->cfg.maxBatchPoints > 0 &&
                        points + r->points > s->cfg.maxBatchPoints)
                        break;
                    points += r->points;
                    s->admitted.pop_front();
                    r->state = RUNNING;
                    r->tLaunch = s->now();
                    batch.push_back(r);
                    sl.reqs.push_back(r->id);
                }
            }
            if (batch.empty()) break;
            const tcv_status st = launch_slot(s, sl, batch);
            if (st != TCV_OK) {
                // nothing of this batch may still be running when its memory goes
                (void)hipStreamSynchronize(sl.ctx->stream);
                std::lock_guard<std::mutex> lk(s->mu);
                s->err = tcv_last_error(sl.ctx);
                for (Req *r : batch) finish(s, *r, FAILED, st);
                sl.reqs.clear();
                sl.alone = false;
                s->cv.notify_all();
                continue;
            }
            sl.busy = true;
            did = true;
        }
        if (!did) {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

void destroy_slots(tcv_server *s)
{
    for (Slot &sl : s->slots) {
        if (sl.ctx) (void)hipStreamSynchronize(sl.ctx->stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        for (tcv_batch *b : {&sl.grid, &sl.degrid}) {
            if (b->d_jobs) (void)hipFree(b->d_jobs);
            if (b->d_prefix) (void)hipFree(b->d_prefix);
            if (b->d_queue) (void)hipFree(b->d_queue);
        }
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        if (sl.ctx) tcv_destroy(sl.ctx);
    }
    s->slots.clear();
}

}  // namespace

extern "C" {

tcv_status tcv_server_create(tcv_ctx *ctx, const tcv_server_config *cfg, tcv_server **out)
{
    if (!ctx || !cfg || !out || cfg->arenaBytes < 4096 || cfg->maxInflight < 1 || cfg->maxInflight > 8 ||
        cfg->maxBatch < 0 || cfg->maxBatchPoints < 0 || cfg->exclusivePoints < 0)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_server_create: bad argument");
    *out = nullptr;
    auto s = std::make_unique<tcv_server>();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->tuning = tcv_tuning{cfg->tile_w, cfg->tile_h, cfg->chunk, cfg->slice_keys, cfg->phases, 0,
                           kDefaultGridKernel, kDefaultDegridKernel, 0, 0, cfg->item_entries,
                           cfg->grid_cols};
    if (s->tuning.tile_w <= 0) s->tuning.tile_w = 64;   // one tile shape for every batch
    if (s->tuning.tile_h <= 0) s->tuning.tile_h = 64;
    if (s->tuning.grid_cols <= 0) s->tuning.grid_cols = auto_grid_cols(s->tuning.tile_w, s->tuning.tile_h);
    TCV_HIP(ctx, hipSetDevice(ctx->device));
    if (hipMalloc((void **)&s->arenaBase, cfg->arenaBytes) != hipSuccess) {
        (void)hipGetLastError();
        return tcv_fail(ctx, TCV_ERR_NOMEM, "tcv_server_create: arena of " +
                                                std::to_string(cfg->arenaBytes) + " bytes failed");
    }
    s->arena.init(cfg->arenaBytes);
    const int maxJobs = cfg->maxBatch > 0 ? cfg->maxBatch : 4096;
    s->slots.resize(cfg->maxInflight);
    tcv_status st = TCV_OK;
    for (Slot &sl : s->slots) {
        if ((st = tcv_create(ctx->device, &sl.ctx)) != TCV_OK) break;
        for (tcv_batch *b : {&sl.grid, &sl.degrid}) {
            b->tw = s->tuning.tile_w;
            b->th = s->tuning.tile_h;
            b->kern = s->tuning.kernel;
            b->dkern = s->tuning.degrid_kernel;
            b->gridCols = s->tuning.grid_cols;
            if (hipMalloc((void **)&b->d_jobs, sizeof(JobDesc) * maxJobs) != hipSuccess ||
                hipMalloc((void **)&b->d_prefix, sizeof(int) * 9 * (maxJobs + 1)) != hipSuccess ||
                hipMalloc((void **)&b->d_queue, kBatchQueueBytes) != hipSuccess)
                st = TCV_ERR_NOMEM;
        }
        if (st == TCV_OK && hipHostMalloc((void **)&sl.pinned, sizeof(JobDesc) * maxJobs * 2) != hipSuccess)
            st = TCV_ERR_NOMEM;
        if (st == TCV_OK && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
            st = TCV_ERR_DEVICE;
        if (st != TCV_OK) break;
    }
    if (st != TCV_OK) {
        destroy_slots(s.get());
        (void)hipFree(s->arenaBase);
        return tcv_fail(ctx, st, "tcv_server_create: slot set-up failed");
    }
    s->t0 = Clock::now();
    tcv_server *raw = s.release();
    raw->th = std::thread(scheduler, raw);
    *out = raw;
    return TCV_OK;
}

tcv_status tcv_server_destroy(tcv_server *s)
{
    if (!s) return TCV_OK;
    s->stop.store(true);
    s->cv.notify_all();
    if (s->th.joinable()) s->th.join();
    destroy_slots(s);
    (void)hipFree(s->arenaBase);
    delete s;
    return TCV_OK;
}

double tcv_server_clock(tcv_server *s) { return s ? s->now() : 0.0; }

// Validates a request and sizes its plan and result (no device work).
static tcv_status prepare(tcv_server *s, const tcv_request *req, Req *r)
{
    if (!req || (req->op != 0 && req->op != 1) || !req->d_C ||
        (req->op == 0 && !req->d_data) || (req->op == 1 && !req->d_grid))
        return TCV_ERR_INVALID;
    r->r = *req;
    const tcv_geometry &G = req->geom;
    if (G.nPlanes < 1 || !G.sSize) return TCV_ERR_INVALID;
    r->sSize.assign(G.sSize, G.sSize + G.nPlanes);
    r->r.geom.sSize = r->sSize.data();
    tcv_status st = plan_setup(s->ctx, &r->r.geom, &s->tuning, &r->plan, &r->planBytes);
    if (st != TCV_OK) return st;
    if (r->plan.sh.tw != s->tuning.tile_w || r->plan.sh.th != s->tuning.tile_h ||
        r->plan.sh.kern != s->tuning.kernel || r->plan.sh.dkern != s->tuning.degrid_kernel)
        return TCV_ERR_UNSUPPORTED;   // every job of a batch runs the same kernels
    r->resBytes = req->op == 0 ? sizeof(float2) * (size_t)G.gSize * G.gSize
                               : sizeof(float2) * (size_t)G.nVis;
    long long pts = 0;
    for (int k = 0; k < G.nPlanes; k++) pts = std::max<long long>(pts, r->sSize[k]);
    r->points = (long long)G.nVis * pts * pts;   // upper bound when widths vary
    return TCV_OK;
}

tcv_status tcv_server_footprint(tcv_server *s, const tcv_request *req, size_t *bytes)
{
    if (!s || !bytes) return TCV_ERR_INVALID;
    Req r;
    tcv_status st = prepare(s, req, &r);
    if (st != TCV_OK) return st;
    *bytes = tcv_detail::Arena::round(r.planBytes) + tcv_detail::Arena::round(r.resBytes);
    return TCV_OK;
}

tcv_status tcv_server_submit(tcv_server *s, const tcv_request *req, long long *id)
{
    if (!s || !id) return TCV_ERR_INVALID;
    auto r = std::make_unique<Req>();
    tcv_status st = prepare(s, req, r.get());
    if (st != TCV_OK) return st;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->arena.fits_ever(tcv_detail::Arena::round(r->planBytes) + r->resBytes)) {
        s->err = "request needs " + std::to_string(r->planBytes + r->resBytes) +
                 " bytes, more than the whole budget";
        return TCV_ERR_BUDGET;
    }
    r->id = s->nextId++;
    r->tSub = s->now();
    *id = r->id;
    s->waiting.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->cv.notify_all();
    return TCV_OK;
}

tcv_status tcv_server_cancel(tcv_server *s, long long id)
{
    if (!s) return TCV_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return TCV_ERR_INVALID;
    Req &r = *it->second;
    if (r.state != QUEUED) return TCV_ERR_UNSUPPORTED;   // already admitted: runs to completion
    finish(s, r, CANCELLED, TCV_ERR_CANCELLED);
    s->cv.notify_all();
    return TCV_OK;
}

tcv_status tcv_server_wait(tcv_server *s, long long id, double timeout_s, tcv_result *res)
{
    if (!s) return TCV_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return TCV_ERR_INVALID;
    Req *r = it->second.get();
    const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(timeout_s));
    while (r->state < DONE) {
        if (s->cv.wait_until(lk, until) == std::cv_status::timeout && r->state < DONE)
            return TCV_ERR_TIMEOUT;   // the request is still pending
    }
    if (res) {
        res->submitted = r->tSub;
        res->admitted = r->tAdm;
        res->launched = r->tLaunch;
        res->completed = r->tDone;
        res->status = r->status;
        res->d_result = r->hasRes ? (const float *)(s->arenaBase + r->resOff) : nullptr;
        res->points = r->points;
    }
    return TCV_OK;
}

tcv_status tcv_server_release(tcv_server *s, long long id)
{
    if (!s) return TCV_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return TCV_ERR_INVALID;
    if (it->second->state < DONE) return TCV_ERR_UNSUPPORTED;
    release_result(s, *it->second);
    s->reqs.erase(it);
    s->cv.notify_all();
    return TCV_OK;
}

tcv_status tcv_server_stats(tcv_server *s, long long *queued, long long *running,
                            size_t *arenaUsed, size_t *arenaCapacity)
{
    if (!s) return TCV_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    long long q = 0, r = 0;
    for (auto &kv : s->reqs) {
        q += kv.second->state == QUEUED;
        r += kv.second->state == ADMITTED || kv.second->state == RUNNING;
    }
    if (queued) *queued = q;
    if (running) *running = r;
    if (arenaUsed) *arenaUsed = s->arena.used();
    if (arenaCapacity) *arenaCapacity = s->arena.capacity();
    return TCV_OK;
}

const char *tcv_server_last_error(tcv_server *s) { return s ? s->err.c_str() : "no server"; }

}  // extern "C"
