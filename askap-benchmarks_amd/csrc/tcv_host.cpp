// Host side of libtcv: context, memory, timing, the Benchmark set-up
// generators, and the C ABI entry points declared in include/tcv.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "tcv_internal.h"

tcv_status tcv_fail(tcv_ctx *ctx, tcv_status s, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return s;
}

ProfScope::ProfScope(tcv_ctx *c, const char *fam) : ctx(c), family(fam)
{
    if (!ctx->profiling) return;
    auto take = [&]() {
        hipEvent_t e = nullptr;
        if (!ctx->pool.empty()) {
            e = ctx->pool.back();
            ctx->pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    };
    a = take();
    b = take();
    if (a) (void)hipEventRecord(a, ctx->stream);
}

ProfScope::~ProfScope()
{
    if (!ctx->profiling || !a || !b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({a, b, family});
}

static tcv_status drain_profile(tcv_ctx *ctx)
{
    for (auto &p : ctx->pending) {
        TCV_HIP(ctx, hipEventSynchronize(p.b));
        float ms = 0.f;
        TCV_HIP(ctx, hipEventElapsedTime(&ms, p.a, p.b));
        auto &st = ctx->stats[p.family];
        st.launches += 1;
        st.ms += ms;
        ctx->pool.push_back(p.a);
        ctx->pool.push_back(p.b);
    }
    ctx->pending.clear();
    return TCV_OK;
}

extern "C" {

const char *tcv_status_string(tcv_status s)
{
    switch (s) {
    case TCV_OK: return "ok";
    case TCV_ERR_INVALID: return "invalid argument";
    case TCV_ERR_RANGE: return "index out of range";
    case TCV_ERR_DEVICE: return "device error";
    case TCV_ERR_NOMEM: return "out of device memory";
    case TCV_ERR_BUDGET: return "HBM budget exceeded";
    case TCV_ERR_UNSUPPORTED: return "unsupported";
    case TCV_ERR_CANCELLED: return "cancelled";
    case TCV_ERR_TIMEOUT: return "timed out";
    }
    return "unknown status";
}

tcv_status tcv_create(int device, tcv_ctx **out)
{
    if (!out) return TCV_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TCV_ERR_DEVICE;
    if (device < 0 || device >= n) return TCV_ERR_INVALID;
    tcv_ctx *ctx = new tcv_ctx();
    ctx->device = device;
    auto fail = [&](tcv_status s) { delete ctx; return s; };
    if (hipSetDevice(device) != hipSuccess) return fail(TCV_ERR_DEVICE);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(TCV_ERR_DEVICE);
    ctx->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(TCV_ERR_DEVICE);
    if (hipEventCreate(&ctx->timer0) != hipSuccess || hipEventCreate(&ctx->timer1) != hipSuccess)
        return fail(TCV_ERR_DEVICE);
    if (hipMalloc(&ctx->d_bcast, kMaxPersistentBlocks * sizeof(int)) != hipSuccess)
        return fail(TCV_ERR_NOMEM);
    if (hipMalloc(&ctx->d_counters, 8 * sizeof(unsigned long long)) != hipSuccess)
        return fail(TCV_ERR_NOMEM);
    if (hipMemset(ctx->d_counters, 0, 8 * sizeof(unsigned long long)) != hipSuccess)
        return fail(TCV_ERR_DEVICE);
    noalloc_setup(ctx);
    *out = ctx;
    return TCV_OK;
}

void tcv_destroy(tcv_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->pool) (void)hipEventDestroy(e);
    if (ctx->timer0) (void)hipEventDestroy(ctx->timer0);
    if (ctx->timer1) (void)hipEventDestroy(ctx->timer1);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_bcast) (void)hipFree(ctx->d_bcast);
    major_release(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *tcv_last_error(const tcv_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int tcv_device_cus(const tcv_ctx *ctx) { return ctx ? ctx->cus : 0; }

tcv_status tcv_synchronize(tcv_ctx *ctx)
{
    if (!ctx) return TCV_ERR_INVALID;
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TCV_OK;
}

tcv_status tcv_malloc(tcv_ctx *ctx, size_t bytes, void **d_ptr)
{
    if (!ctx || !d_ptr) return TCV_ERR_INVALID;
    *d_ptr = nullptr;
    TCV_HIP(ctx, hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(d_ptr, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return tcv_fail(ctx, TCV_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return TCV_OK;
}

tcv_status tcv_mem_info(tcv_ctx *ctx, size_t *freeBytes, size_t *totalBytes)
{
    if (!ctx || !freeBytes || !totalBytes) return TCV_ERR_INVALID;
    TCV_HIP(ctx, hipSetDevice(ctx->device));
    TCV_HIP(ctx, hipMemGetInfo(freeBytes, totalBytes));
    return TCV_OK;
}

tcv_status tcv_free(tcv_ctx *ctx, void *d_ptr)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!d_ptr) return TCV_OK;
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    TCV_HIP(ctx, hipFree(d_ptr));
    return TCV_OK;
}

tcv_status tcv_upload(tcv_ctx *ctx, void *d_dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!d_dst || !src))) return TCV_ERR_INVALID;
    if (!bytes) return TCV_OK;
    TCV_HIP(ctx, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TCV_OK;
}

tcv_status tcv_download(tcv_ctx *ctx, void *dst, const void *d_src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !d_src))) return TCV_ERR_INVALID;
    if (!bytes) return TCV_OK;
    TCV_HIP(ctx, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TCV_OK;
}

tcv_status tcv_zero(tcv_ctx *ctx, void *d_ptr, size_t bytes)
{
    if (!ctx || (bytes && !d_ptr) || (bytes % 4)) return TCV_ERR_INVALID;
    return launch_fill_u32(ctx, d_ptr, 0u, bytes / 4);
}

tcv_status tcv_copy(tcv_ctx *ctx, void *d_dst, const void *d_src, size_t bytes)
{
    if (!ctx) return TCV_ERR_INVALID;
    if ((bytes % 16) || (bytes && (!d_dst || !d_src)) || ((uintptr_t)d_dst % 16) ||
        ((uintptr_t)d_src % 16))
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_copy: pointers and size must be 16-byte multiples");
    return launch_copy(ctx, d_dst, d_src, bytes / 16);
}

tcv_status tcv_read_bandwidth(tcv_ctx *ctx, const void *d_buf, size_t bytes, int rounds,
                              double *gbps)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!d_buf || !gbps || rounds < 8 || (rounds % 8) || bytes < 4096 || (bytes & (bytes - 1)) ||
        bytes > (64ull << 30) || ((uintptr_t)d_buf % 16))
        return tcv_fail(ctx, TCV_ERR_INVALID,
                        "tcv_read_bandwidth: 16-byte aligned power-of-two buffer of 4 KiB..64 GiB, "
                        "rounds a positive multiple of 8");
    if (!ctx->d_bcast) return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_read_bandwidth: context not set up");
    unsigned *sink = reinterpret_cast<unsigned *>(ctx->d_bcast);   // written only by chance
    tcv_status st = launch_read_stream(ctx, d_buf, bytes / 16, rounds, sink);   // warm-up
    if (st != TCV_OK) return st;
    const int reps = 5;
    TCV_HIP(ctx, hipEventRecord(ctx->timer0, ctx->stream));
    for (int r = 0; r < reps && st == TCV_OK; ++r) st = launch_read_stream(ctx, d_buf, bytes / 16, rounds, sink);
    if (st != TCV_OK) return st;
    TCV_HIP(ctx, hipEventRecord(ctx->timer1, ctx->stream));
    TCV_HIP(ctx, hipEventSynchronize(ctx->timer1));
    float ms = 0.f;
    TCV_HIP(ctx, hipEventElapsedTime(&ms, ctx->timer0, ctx->timer1));
    const double window = (double)std::max(1, ctx->cus) * 8.0 * 256.0 * 16.0;   // bytes per round
    *gbps = reps * window * rounds / (ms * 1e-3) / 1e9;
    return TCV_OK;
}

tcv_status tcv_timer_start(tcv_ctx *ctx)
{
    if (!ctx) return TCV_ERR_INVALID;
    TCV_HIP(ctx, hipEventRecord(ctx->timer0, ctx->stream));
    return TCV_OK;
}

tcv_status tcv_timer_stop(tcv_ctx *ctx, double *ms)
{
    if (!ctx || !ms) return TCV_ERR_INVALID;
    TCV_HIP(ctx, hipEventRecord(ctx->timer1, ctx->stream));
    TCV_HIP(ctx, hipEventSynchronize(ctx->timer1));
    float f = 0.f;
    TCV_HIP(ctx, hipEventElapsedTime(&f, ctx->timer0, ctx->timer1));
    *ms = f;
    return TCV_OK;
}

tcv_status tcv_set_profiling(tcv_ctx *ctx, int on)
{
    if (!ctx) return TCV_ERR_INVALID;
    ctx->profiling = on != 0;
    return TCV_OK;
}

tcv_status tcv_kernel_stats(tcv_ctx *ctx, const char *family, long *launches, double *total_ms)
{
    if (!ctx || !family) return TCV_ERR_INVALID;
    tcv_status st = drain_profile(ctx);
    if (st != TCV_OK) return st;
    auto it = ctx->stats.find(family);
    if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.ms;
    return TCV_OK;
}

tcv_status tcv_reset_stats(tcv_ctx *ctx)
{
    if (!ctx) return TCV_ERR_INVALID;
    tcv_status st = drain_profile(ctx);
    ctx->stats.clear();
    return st;
}

// ---- a6 ----------------------------------------------------------------------

tcv_status tcv_init_offsets(tcv_ctx *ctx, const double *d_u, const double *d_v, const double *d_w,
                            int nSamples, const double *freq, int nChan, double cellSize,
                            double wCellSize, int wSize, int gSize, int overSample,
                            const int *sSize, const int *cOffset0, int *d_iu, int *d_iv,
                            int *d_wPlane, int *d_cOffset, long long *numPerPlane,
                            long long *nOutOfRange)
{
    if (!ctx || nSamples < 0 || nChan < 1 || !freq || wSize < 1 || gSize < 1 || !sSize ||
        overSample < 1 || cellSize <= 0 || (nSamples > 0 && (!d_u || !d_v || !d_w || !d_iu || !d_iv || !d_cOffset)))
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_init_offsets: bad argument");
    if ((long long)nSamples * nChan > 0x7fffffffLL)
        return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "more than 2^31-1 visibilities");
    if (cOffset0 == nullptr && wCellSize <= 0)
        return tcv_fail(ctx, TCV_ERR_INVALID, "attic layout needs wCellSize > 0");
    TCV_HIP(ctx, hipSetDevice(ctx->device));
    const int nS = cOffset0 ? wSize : 1;
    // small staging block: freq | sSize | cOffset0 | numPerPlane | bad
    const size_t fB = sizeof(double) * nChan, sB = sizeof(int) * nS, oB = sizeof(int) * wSize;
    const size_t nB = sizeof(unsigned long long) * wSize;
    const size_t total = fB + sB + oB + nB + 64;
    char *d = nullptr;
    TCV_HIP(ctx, hipMalloc((void **)&d, total + 64));
    double *d_freq = (double *)d;
    int *d_s = (int *)(d + fB);
    int *d_o = (int *)(d + fB + sB);
    size_t nOff = (fB + sB + oB + 7) & ~size_t(7);
    unsigned long long *d_npp = (unsigned long long *)(d + nOff);
    unsigned long long *d_bad = d_npp + wSize;
    auto cleanup = [&]() { (void)hipFree(d); };
    hipError_t e = hipMemcpyAsync(d_freq, freq, fB, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_s, sSize, sB, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && cOffset0)
        e = hipMemcpyAsync(d_o, cOffset0, oB, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        cleanup();
        return tcv_fail(ctx, TCV_ERR_DEVICE, hipGetErrorString(e));
    }
    tcv_status st = launch_fill_u32(ctx, d_npp, 0, (nB + 8) / 4);
    if (st == TCV_OK)
        st = launch_offsets(ctx, d_u, d_v, d_w, nSamples, d_freq, nChan, cellSize, wCellSize,
                            wSize, gSize, overSample, d_s, cOffset0 ? d_o : nullptr, d_iu, d_iv,
                            d_wPlane, d_cOffset, d_npp, d_bad);
    std::vector<unsigned long long> npp(wSize + 1);
    if (st == TCV_OK) {
        e = hipMemcpyAsync(npp.data(), d_npp, nB + 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = tcv_fail(ctx, TCV_ERR_DEVICE, hipGetErrorString(e));
    }
    cleanup();
    if (st != TCV_OK) return st;
    if (numPerPlane)
        for (int k = 0; k < wSize; k++) numPerPlane[k] = (long long)npp[k];
    const long long bad = (long long)npp[wSize];
    if (nOutOfRange) *nOutOfRange = bad;
    if (bad) return tcv_fail(ctx, TCV_ERR_RANGE, std::to_string(bad) + " visibilities out of range");
    return TCV_OK;
}

}  // extern "C"

// ---- geometry helpers ------------------------------------------------------------

static tcv_status check_geometry(tcv_ctx *ctx, const tcv_geometry *g, int *maxS)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!g || g->nVis < 0 || g->gSize < 1 || g->nPlanes < 1 || !g->sSize || g->cSize < 0)
        return tcv_fail(ctx, TCV_ERR_INVALID, "bad geometry");
    if (g->nVis > 0 && (!g->d_iu || !g->d_iv || !g->d_cOffset))
        return tcv_fail(ctx, TCV_ERR_INVALID, "geometry index arrays missing");
    if (!g->d_wPlane && g->nPlanes != 1)
        return tcv_fail(ctx, TCV_ERR_INVALID, "uniform geometry needs nPlanes == 1");
    int m = 0;
    for (int k = 0; k < g->nPlanes; k++) {
        if (g->sSize[k] < 1) return tcv_fail(ctx, TCV_ERR_INVALID, "kernel width < 1");
        m = std::max(m, g->sSize[k]);
    }
    *maxS = m;
    return TCV_OK;
}

// Uploads the per-plane widths into a small device buffer when needed.
struct GeomUpload {
    int *d_s = nullptr;
    ~GeomUpload() { if (d_s) (void)hipFree(d_s); }
};

static tcv_status make_geom_dev(tcv_ctx *ctx, const tcv_geometry *g, GeomUpload &up, GeomDev &gd)
{
    gd.nVis = g->nVis;
    gd.iu = g->d_iu;
    gd.iv = g->d_iv;
    gd.cOffset = g->d_cOffset;
    gd.wPlane = g->d_wPlane;
    gd.nPlanes = g->nPlanes;
    gd.sUniform = g->sSize[0];
    gd.gSize = g->gSize;
    gd.cSize = g->cSize;
    gd.sSize = nullptr;
    if (g->d_wPlane) {
        TCV_HIP(ctx, hipMalloc((void **)&up.d_s, sizeof(int) * g->nPlanes));
        TCV_HIP(ctx, hipMemcpyAsync(up.d_s, g->sSize, sizeof(int) * g->nPlanes,
                                    hipMemcpyHostToDevice, ctx->stream));
        gd.sSize = up.d_s;
    }
    return TCV_OK;
}

extern "C" {

tcv_status tcv_grid_plain(tcv_ctx *ctx, const tcv_geometry *g, const float *d_data,
                          const float *d_C, float *d_grid)
{
    int maxS = 0;
    tcv_status st = check_geometry(ctx, g, &maxS);
    if (st != TCV_OK) return st;
    if (g->nVis > 0 && (!d_data || !d_C || !d_grid))
        return tcv_fail(ctx, TCV_ERR_INVALID, "null data, C or grid");
    GeomUpload up;
    GeomDev gd;
    if ((st = make_geom_dev(ctx, g, up, gd)) != TCV_OK) return st;
    st = launch_grid_plain(ctx, gd, maxS, (const float2 *)d_data, (const float2 *)d_C,
                           (float2 *)d_grid);
    if (up.d_s) TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st;
}

tcv_status tcv_degrid_plain(tcv_ctx *ctx, const tcv_geometry *g, const float *d_grid,
                            const float *d_C, float *d_data)
{
    int maxS = 0;
    tcv_status st = check_geometry(ctx, g, &maxS);
    if (st != TCV_OK) return st;
    if (g->nVis > 0 && (!d_data || !d_C || !d_grid))
        return tcv_fail(ctx, TCV_ERR_INVALID, "null data, C or grid");
    GeomUpload up;
    GeomDev gd;
    if ((st = make_geom_dev(ctx, g, up, gd)) != TCV_OK) return st;
    st = launch_degrid_plain(ctx, gd, (const float2 *)d_grid, (const float2 *)d_C,
                             (float2 *)d_data);
    if (up.d_s) TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st;
}

tcv_status tcv_grid_plain_f64(tcv_ctx *ctx, const tcv_geometry *g, const float *d_data,
                              const float *d_C, double *d_grid)
{
    int maxS = 0;
    tcv_status st = check_geometry(ctx, g, &maxS);
    if (st != TCV_OK) return st;
    if (g->nVis > 0 && (!d_C || !d_grid)) return tcv_fail(ctx, TCV_ERR_INVALID, "null C or grid");
    GeomUpload up;
    GeomDev gd;
    if ((st = make_geom_dev(ctx, g, up, gd)) != TCV_OK) return st;
    st = launch_grid_plain_f64(ctx, gd, (const float2 *)d_data, (const float2 *)d_C, (double2 *)d_grid);
    if (up.d_s) TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st;
}

tcv_status tcv_degrid_plain_f64(tcv_ctx *ctx, const tcv_geometry *g, const float *d_grid,
                                const float *d_C, double *d_data)
{
    int maxS = 0;
    tcv_status st = check_geometry(ctx, g, &maxS);
    if (st != TCV_OK) return st;
    if (g->nVis > 0 && (!d_data || !d_C || !d_grid))
        return tcv_fail(ctx, TCV_ERR_INVALID, "null data, C or grid");
    GeomUpload up;
    GeomDev gd;
    if ((st = make_geom_dev(ctx, g, up, gd)) != TCV_OK) return st;
    st = launch_degrid_plain_f64(ctx, gd, (const float2 *)d_grid, (const float2 *)d_C, (double2 *)d_data);
    if (up.d_s) TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st;
}

tcv_status tcv_range_errors(tcv_ctx *ctx, long long *count)
{
    if (!ctx || !count) return TCV_ERR_INVALID;
    unsigned long long c = 0;
    TCV_HIP(ctx, hipMemcpyAsync(&c, ctx->d_counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = (long long)c;
    return launch_fill_u32(ctx, ctx->d_counters, 0, 2);
}

// ---- plans -------------------------------------------------------------------

// Defaults (a tuning field <= 0 means auto; xcd_split < 0 means auto):
// tile width from the kernel width, 64-row tiles, 2048-entry work items, up to
// 2048 C-slice keys (the binning's (tile, key) histogram stays small enough for
// its atomics to be cheap; finer keys ordered the entries no better for the
// kernels), entries ordered by key in phases of about 5.5 MB of taps
// (concurrently running workgroups share the phase's slices in L2; fewer
// phases when work items would get short), and, with 8 or more phases, work
// items ordered XCD-major with one queue per XCD so each XCD walks its own
// phases.  Kernels: kDefaultGridKernel / kDefaultDegridKernel.
// C-slice phases: about `bytesPerPhase` of taps per phase, so that the slices
// one phase touches stay in L2 / Infinity Cache while its tiles run (measured
// best: ~5.5 MB for both kernels), but no more than leaves about 64 entries
// per (tile, phase) work item on average: every item zeroes and flushes its
// tile.  From 8 phases up a multiple of 8, so the phases split evenly over
// the XCDs.
static int default_phases(long long cSize, int tapBytes, int sliceKeys, const tcv_geometry *g, int maxS,
                          int tw, int th, double bytesPerPhase, double itemEntries)
{
    double target = std::max(1.0, (double)cSize * tapBytes / bytesPerPhase);
    if (g && tw > 0 && th > 0) {
        const double perVis = (1.0 + (maxS - 1.0) / tw) * (1.0 + (maxS - 1.0) / th);
        const double tiles = (double)((g->gSize + tw - 1) / tw) * ((g->gSize + th - 1) / th);
        target = std::min(target, std::max(1.0, g->nVis * perVis / tiles / itemEntries));
    }
    target = std::min(target, (double)sliceKeys);
    if (target >= 8) {
        const int ph = (int)(target / 8.0 + 0.5) * 8;   // nearest multiple of 8
        return std::min(ph, sliceKeys / 8 * 8);
    }
    int ph = 1;
    while (ph * 2 <= target) ph *= 2;
    return ph;
}

constexpr double kDegridPhaseBytes = 5.5e6;
constexpr int kItemEntries = 64;   // entries per work item the phase count keeps at least
constexpr double kGridPhaseBytes = 5.5e6;   // the gridder may use its own count (grid_phases)
// 128 x 64 tiles carry half again as many entries per work item: smaller phases
// (measured optimum 256-320 phases on a 1.1 GB table, against 184-216 for 64 x 64)
constexpr double kWidePhaseBytes = 3.8e6;
constexpr int kWideTileMinWidth = 89;

static void auto_tuning(int maxS, const tcv_geometry *g, tcv_tuning *t)
{
    int minS = maxS;
    for (int k = 0; k < g->nPlanes; k++) minS = std::min(minS, g->sSize[k]);
    // 128-wide tiles (pair kernels only) for kernels 89 wide and up: a footprint
    // row meets fewer tiles, each visibility makes fewer entries, and the
    // degridder's eight-wave workgroups walk them.  Grid + degrid measured 20 %
    // faster at width 129, 10-11 % at 113 and 121, 6 % at 97, equal at 81 and
    // 4 % slower at 65.  (FP16 storage keeps 64-wide tiles.)
    if (t->tap_bytes <= 0) t->tap_bytes = 8;
    if (t->item_entries <= 0) t->item_entries = kItemEntries;
    const bool pairs = minS >= 2 && (t->kernel <= 0 || t->kernel == 3) &&
                       (t->degrid_kernel <= 0 || t->degrid_kernel == 3);
    if (t->tile_w <= 0)
        t->tile_w = (maxS >= kWideTileMinWidth && pairs && t->tile_h <= 64 && t->tap_bytes == 8) ? 128
                    : maxS >= 48 ? 64 : maxS >= 24 ? 32 : maxS >= 12 ? 16 : 8;
    if (t->tile_h <= 0) t->tile_h = 64;
    if (t->grid_cols <= 0) t->grid_cols = auto_grid_cols(t->tile_w, t->tile_h);
    if (t->chunk <= 0)
        // narrow kernels make short entries, and shorter work items measured
        // faster for them (more items to share out, none running long at the
        // end): 512 for jobs of narrow kernels only (current runType 1, widths
        // 7-19: grid 31 % and degrid 18 % faster than 2048), 1024 for jobs with
        // some (runType 0, widths 7-121: 10 % and 3 %); wide kernels keep 2048
        // (config 1's degridder is 6 % slower at 512)
        t->chunk = maxS < 24 ? 512 : minS < 24 ? 1024 : 2048;
    const long long slices = std::max<long long>(1, g->cSize / ((long long)minS * minS));
    if (t->slice_keys <= 0) t->slice_keys = (int)std::min<long long>(2048, slices);
    // 8192-cell tiles (128 x 64, 64 x 128, 32 x 256) carry more entries per item
    const bool wide = t->tile_w * t->tile_h > 4096;
    if (t->phases <= 0)
        t->phases = default_phases(g->cSize, t->tap_bytes, t->slice_keys, g, maxS, t->tile_w, t->tile_h,
                                   wide ? kWidePhaseBytes : kDegridPhaseBytes, t->item_entries);
    if (t->xcd_split < 0) t->xcd_split = (t->phases >= 8 && t->phases % 8 == 0) ? 1 : 0;
    if (t->grid_phases <= 0)
        t->grid_phases = default_phases(g->cSize, t->tap_bytes, t->slice_keys, g, maxS, t->tile_w,
                                        t->tile_h, wide ? kWidePhaseBytes : kGridPhaseBytes, t->item_entries);
    if (t->kernel <= 0) t->kernel = kDefaultGridKernel;
    if (t->degrid_kernel <= 0) t->degrid_kernel = kDefaultDegridKernel;
}

static void free_plan(tcv_plan *p)
{
    if (p->ownsBlock && p->block) (void)hipFree(p->block);
    p->block = nullptr;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // extern "C"

tcv_status plan_setup(tcv_ctx *ctx, const tcv_geometry *g, const tcv_tuning *tuning, tcv_plan *p,
                      size_t *bytes)
{
    int maxS = 0;
    tcv_status st = check_geometry(ctx, g, &maxS);
    if (st != TCV_OK) return st;
    if (g->gSize > 65535) return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tiled path needs gSize <= 65535");
    if (g->cSize * 8 >= 0x7fffffffLL)
        return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tiled path needs a C table below 2 GiB");
    tcv_tuning t = tuning ? *tuning : tcv_tuning{0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0};
    if (t.tap_bytes != 0 && t.tap_bytes != 4 && t.tap_bytes != 8)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tuning.tap_bytes must be 0, 4 or 8");
    if (t.item_entries < 0) return tcv_fail(ctx, TCV_ERR_INVALID, "tuning.item_entries must be >= 0");
    if (t.grid_cols < 0) return tcv_fail(ctx, TCV_ERR_INVALID, "tuning.grid_cols must be >= 0");
    if (!tuning && ctx) {
        auto it = ctx->tuned.find(tuning_class(g));
        if (it != ctx->tuned.end()) t = it->second;
    }
    auto_tuning(maxS, g, &t);
    if (!tiled_config_supported(t.tile_w, t.tile_h))
        return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tile shape not built");
    if (t.chunk < 1 || t.phases < 1 || t.slice_keys < 1)
        return tcv_fail(ctx, TCV_ERR_INVALID, "bad tuning");
    if (!grid_cols_supported(t.tile_w, t.tile_h, t.grid_cols))
        return tcv_fail(ctx, TCV_ERR_UNSUPPORTED,
                        "tuning.grid_cols must be tile_w, or tile_w / 2 or / 4 (at least 16) on tiles 64 wide "
                        "and up with tile_h >= 512 / grid_cols");
    PlanShape &sh = p->sh;
    sh.tw = t.tile_w;
    sh.th = t.tile_h;
    sh.chunk = t.chunk;
    sh.k2 = std::max(1, t.slice_keys);
    sh.phases = std::min(t.phases, sh.k2);
    sh.xcd = (t.xcd_split > 0 && sh.phases % 8 == 0) ? 1 : 0;
    sh.phasesG = std::min(t.grid_phases, sh.k2);
    sh.xcdG = (t.xcd_split > 0 && sh.phasesG % 8 == 0) ? 1 : 0;
    if (t.kernel < 1 || t.kernel > 4 || t.degrid_kernel < 1 || t.degrid_kernel > 4)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tuning.kernel and degrid_kernel must be 0..4");
    sh.kern = t.kernel;
    sh.dkern = t.degrid_kernel;
    sh.tapBytes = t.tap_bytes;
    sh.itemEntries = t.item_entries;
    sh.gridCols = t.grid_cols;
    // the pair kernels need kernel widths of at least 2 (a pair inside each C row)
    int minS = maxS;
    for (int k = 0; k < g->nPlanes; k++) minS = std::min(minS, g->sSize[k]);
    if (minS < 2 || (sh.kern >= 3 && !grid2_supported(sh.tw, sh.th))) {
        if (sh.kern >= 3) sh.kern = 1;
    }
    if (minS < 2 && sh.dkern >= 3) sh.dkern = 2;
    if ((sh.tw > 64 || sh.th > 64) && (sh.kern != 3 || sh.dkern != 3))
        return tcv_fail(ctx, TCV_ERR_UNSUPPORTED,
                        "tiles wider or taller than 64 need the pair kernels (code 3) and kernel widths "
                        "of at least 2");
    sh.minS = minS;
    sh.nTX = (g->gSize + sh.tw - 1) / sh.tw;
    sh.nTY = (g->gSize + sh.th - 1) / sh.th;
    sh.nTiles = sh.nTX * sh.nTY;
    sh.maxS = maxS;
    const long long perVis = (long long)((maxS + sh.tw - 2) / sh.tw + 1) * ((maxS + sh.th - 2) / sh.th + 1);
    sh.maxEntries = std::max<long long>(1, (long long)g->nVis * perVis);
    sh.nKeys = (long long)sh.nTiles * sh.k2;
    if (sh.maxEntries >= 0x7fffffffLL || sh.nKeys >= 0x7fffffffLL)
        return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "plan too large for 32-bit entry indices");
    p->maxItems = sh.maxEntries / sh.chunk + (long long)sh.phases * sh.nTiles + 1;
    const bool twoLists = sh.phasesG != sh.phases || sh.xcdG != sh.xcd;
    p->maxItemsG = twoLists ? sh.maxEntries / sh.chunk + (long long)sh.phasesG * sh.nTiles + 1 : 0;
    const long long slotsG = (long long)sh.phasesG * sh.nTiles;
    p->nVis = g->nVis;
    p->gSize = g->gSize;
    p->nPlanes = g->nPlanes;
    p->cSize = g->cSize;
    p->sSizeHost.assign(g->sSize, g->sSize + g->nPlanes);
    const long long slots = (long long)sh.phases * sh.nTiles;
    p->scanTmpLen = std::max(scan_tmp_len(sh.nKeys), std::max(scan_tmp_len(slots), scan_tmp_len(slotsG)));
    size_t n = 0;
    n += align256(sizeof(int) * g->nPlanes);
    n += align256(sizeof(int4) * sh.maxEntries);
    if (plan_two_level(sh)) n += align256(sizeof(int4) * sh.maxEntries);
    n += align256(sizeof(int4) * p->maxItems);
    n += 3 * align256(sizeof(int) * (sh.nKeys + 1));
    n += 2 * align256(sizeof(int) * (slots + 1));
    n += align256(sizeof(int) * 32);
    n += align256(sizeof(int) * p->scanTmpLen);
    if (twoLists) {
        n += align256(sizeof(int4) * p->maxItemsG);
        n += 2 * align256(sizeof(int) * (slotsG + 1));
        n += align256(sizeof(int) * 32);
    }
    *bytes = n;
    return TCV_OK;
}

void plan_bind(tcv_plan *p, char *block)
{
    const PlanShape &sh = p->sh;
    const long long slots = (long long)sh.phases * sh.nTiles;
    char *q = block;
    auto take = [&](size_t bytes) {
        char *r = q;
        q += align256(bytes);
        return r;
    };
    p->d_sSize = (int *)take(sizeof(int) * p->nPlanes);
    p->entries = (int4 *)take(sizeof(int4) * sh.maxEntries);
    p->entriesTmp = plan_two_level(sh) ? (int4 *)take(sizeof(int4) * sh.maxEntries) : nullptr;
    p->items = (int4 *)take(sizeof(int4) * p->maxItems);
    p->counts = (int *)take(sizeof(int) * (sh.nKeys + 1));
    p->starts = (int *)take(sizeof(int) * (sh.nKeys + 1));
    p->cursor = (int *)take(sizeof(int) * (sh.nKeys + 1));
    p->itemCount = (int *)take(sizeof(int) * (slots + 1));
    p->itemStart = (int *)take(sizeof(int) * (slots + 1));
    p->meta = (int *)take(sizeof(int) * 32);
    p->scanTmp = (int *)take(sizeof(int) * p->scanTmpLen);
    if (p->maxItemsG) {
        const long long slotsG = (long long)sh.phasesG * sh.nTiles;
        p->itemsG = (int4 *)take(sizeof(int4) * p->maxItemsG);
        p->itemCountG = (int *)take(sizeof(int) * (slotsG + 1));
        p->itemStartG = (int *)take(sizeof(int) * (slotsG + 1));
        p->metaG = (int *)take(sizeof(int) * 32);
    } else {
        p->itemsG = p->items;
        p->itemCountG = p->itemCount;
        p->itemStartG = p->itemStart;
        p->metaG = p->meta;
    }
    p->block = block;
}

void mark_shared_grids(std::vector<JobDesc> &jobs)
{
    // grid byte ranges [grid, grid + 8 gSize^2); every pair compared (batches
    // hold tens to hundreds of jobs)
    const size_t n = jobs.size();
    std::vector<uintptr_t> b(n), e(n);
    for (size_t j = 0; j < n; ++j) {
        jobs[j].sharedGrid = 0;
        b[j] = reinterpret_cast<uintptr_t>(jobs[j].grid);
        e[j] = b[j] + sizeof(float2) * (uintptr_t)jobs[j].gSize * (uintptr_t)jobs[j].gSize;
    }
    for (size_t j = 0; j < n; ++j)
        for (size_t k = j + 1; k < n; ++k)
            if (jobs[j].grid && jobs[k].grid && b[j] < e[k] && b[k] < e[j])
                jobs[j].sharedGrid = jobs[k].sharedGrid = 1;
}

extern "C" {

tcv_status tcv_plan_create(tcv_ctx *ctx, const tcv_geometry *g, const tcv_tuning *tuning,
                           tcv_plan **out)
{
    if (!out) return TCV_ERR_INVALID;
    *out = nullptr;
    if (!ctx) return TCV_ERR_INVALID;
    TCV_HIP(ctx, hipSetDevice(ctx->device));
    tcv_plan *p = new tcv_plan();
    size_t bytes = 0;
    tcv_status st = plan_setup(ctx, g, tuning, p, &bytes);
    if (st != TCV_OK) {
        delete p;
        return st;
    }
    void *block = nullptr;
    if (hipMalloc(&block, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete p;
        return tcv_fail(ctx, TCV_ERR_NOMEM, "plan allocation of " + std::to_string(bytes) + " bytes failed");
    }
    plan_bind(p, (char *)block);
    p->ownsBlock = true;
    if (hipMemcpy(p->d_sSize, g->sSize, sizeof(int) * g->nPlanes, hipMemcpyHostToDevice) != hipSuccess) {
        free_plan(p);
        delete p;
        return tcv_fail(ctx, TCV_ERR_DEVICE, "plan upload failed");
    }
    st = tcv_plan_build(ctx, p, g);
    if (st == TCV_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = tcv_fail(ctx, TCV_ERR_DEVICE, hipGetErrorString(e));
    }
    if (st != TCV_OK) {
        free_plan(p);
        delete p;
        return st;
    }
    *out = p;
    return TCV_OK;
}

tcv_status tcv_plan_build(tcv_ctx *ctx, tcv_plan *p, const tcv_geometry *g)
{
    int maxS = 0;
    tcv_status st = check_geometry(ctx, g, &maxS);
    if (st != TCV_OK) return st;
    if (!p) return tcv_fail(ctx, TCV_ERR_INVALID, "null plan");
    if (g->nVis > p->nVis || g->gSize != p->gSize || g->nPlanes != p->nPlanes || maxS > p->sh.maxS ||
        g->cSize > p->cSize ||
        !std::equal(g->sSize, g->sSize + g->nPlanes, p->sSizeHost.begin()))
        return tcv_fail(ctx, TCV_ERR_INVALID, "geometry does not match the plan");
    GeomDev gd;
    gd.nVis = g->nVis;
    gd.iu = g->d_iu;
    gd.iv = g->d_iv;
    gd.cOffset = g->d_cOffset;
    gd.wPlane = g->d_wPlane;
    gd.nPlanes = g->nPlanes;
    gd.sUniform = g->sSize[0];
    gd.gSize = g->gSize;
    gd.cSize = g->cSize;
    gd.sSize = g->d_wPlane ? p->d_sSize : nullptr;
    p->cSize = g->cSize;
    p->geom = gd;
    p->hasGeom = true;
    return plan_build_kernels(ctx, p, gd);
}

tcv_status tcv_plan_destroy(tcv_ctx *ctx, tcv_plan *p)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!p) return TCV_OK;
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_plan(p);
    delete p;
    return TCV_OK;
}

tcv_status tcv_plan_info(tcv_ctx *ctx, tcv_plan *p, tcv_tuning *t, long long *entries,
                         long long *items, long long *outOfRange)
{
    if (!ctx || !p) return TCV_ERR_INVALID;
    if (t) *t = tcv_tuning{p->sh.tw, p->sh.th, p->sh.chunk, p->sh.k2, p->sh.phases, p->sh.xcd, p->sh.kern,
                           p->sh.dkern, p->sh.phasesG, p->sh.tapBytes, p->sh.itemEntries,
                           p->sh.gridCols};
    int meta[4] = {0, 0, 0, 0};   // items, queue head, entries, out of range
    TCV_HIP(ctx, hipMemcpyAsync(meta, p->meta, sizeof(meta), hipMemcpyDeviceToHost, ctx->stream));
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (items) *items = meta[0];
    if (entries) *entries = meta[2];
    if (outOfRange) *outOfRange = meta[3];
    return TCV_OK;
}

tcv_status tcv_grid_tiled(tcv_ctx *ctx, tcv_plan *p, const float *d_data, const float *d_C,
                          float *d_grid)
{
    if (!ctx || !p || !d_data || !d_C || !d_grid) return tcv_fail(ctx, TCV_ERR_INVALID, "null argument");
    return launch_grid_tiled(ctx, p, d_data, d_C, (float2 *)d_grid, nullptr);
}

tcv_status tcv_grid_psf(tcv_ctx *ctx, tcv_plan *p, const float *d_C, float *d_grid)
{
    if (!ctx || !p || !d_C || !d_grid) return tcv_fail(ctx, TCV_ERR_INVALID, "null argument");
    return launch_grid_tiled(ctx, p, nullptr, d_C, (float2 *)d_grid, nullptr);
}

tcv_status tcv_degrid_tiled(tcv_ctx *ctx, tcv_plan *p, const float *d_grid, const float *d_C,
                            float *d_data)
{
    if (!ctx || !p || !d_data || !d_C || !d_grid) return tcv_fail(ctx, TCV_ERR_INVALID, "null argument");
    return launch_degrid_tiled(ctx, p, (const float2 *)d_grid, d_C, (float2 *)d_data, nullptr);
}

// ---- FP16 storage mode ---------------------------------------------------------

tcv_status tcv_ctable_to_half(tcv_ctx *ctx, const float *d_C, long long nTaps, int nPlanes,
                              const long long *planeStart, long long planeTaps, void *d_Chalf,
                              float *d_planeScale, float *planeScale)
{
    if (!ctx || nTaps < 0 || nPlanes < 1 || (nTaps > 0 && (!d_C || !d_Chalf || !d_planeScale)) ||
        (!planeStart && planeTaps <= 0))
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_ctable_to_half: bad argument");
    if (!planeStart && planeTaps * nPlanes < nTaps)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_ctable_to_half: planes do not cover the table");
    long long *d_ps = nullptr;
    if (planeStart) {
        if (planeStart[0] != 0 || planeStart[nPlanes] != nTaps)
            return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_ctable_to_half: planeStart must span [0, nTaps]");
        for (int k = 0; k < nPlanes; k++)
            if (planeStart[k + 1] < planeStart[k])
                return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_ctable_to_half: planeStart not sorted");
        TCV_HIP(ctx, hipMalloc((void **)&d_ps, sizeof(long long) * (nPlanes + 1)));
        hipError_t e = hipMemcpyAsync(d_ps, planeStart, sizeof(long long) * (nPlanes + 1),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(d_ps);
            return tcv_fail(ctx, TCV_ERR_DEVICE, hipGetErrorString(e));
        }
    }
    // the scale array doubles as the max-abs accumulator (float bit patterns)
    tcv_status st = launch_ctable_to_half(ctx, (const float2 *)d_C, nTaps, nPlanes, d_ps,
                                          planeStart ? 0 : planeTaps, (unsigned *)d_planeScale,
                                          d_Chalf);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (d_ps) (void)hipFree(d_ps);
    if (st != TCV_OK) return st;
    if (e != hipSuccess) return tcv_fail(ctx, TCV_ERR_DEVICE, hipGetErrorString(e));
    if (planeScale)
        TCV_HIP(ctx, hipMemcpy(planeScale, d_planeScale, sizeof(float) * nPlanes, hipMemcpyDeviceToHost));
    return TCV_OK;
}

tcv_status tcv_vis_to_half(tcv_ctx *ctx, const float *d_data, int nVis, void *d_dhalf, float *scale)
{
    if (!ctx || nVis < 0 || !scale || (nVis > 0 && (!d_data || !d_dhalf)))
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_vis_to_half: bad argument");
    unsigned *d_max = reinterpret_cast<unsigned *>(ctx->d_counters + 2);
    tcv_status st = launch_vis_to_half(ctx, (const float2 *)d_data, nVis, d_max, d_dhalf);
    if (st != TCV_OK) return st;
    unsigned bits = 0;
    TCV_HIP(ctx, hipMemcpyAsync(&bits, d_max, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream));
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float f;
    std::memcpy(&f, &bits, sizeof(f));
    *scale = f > 0.f ? f : 1.f;
    return TCV_OK;
}

tcv_status tcv_grid_tiled_half(tcv_ctx *ctx, tcv_plan *p, const void *d_dhalf, float dataScale,
                               const void *d_Chalf, const float *d_planeScale, long long planeTaps,
                               float *d_grid)
{
    if (!ctx || !p || !d_dhalf || !d_Chalf || !d_planeScale || !d_grid || planeTaps < 0 ||
        planeTaps > 0x7fffffffLL)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_grid_tiled_half: bad argument");
    const HalfStorage h{dataScale, d_planeScale, (int)planeTaps};
    return launch_grid_tiled(ctx, p, d_dhalf, d_Chalf, (float2 *)d_grid, &h);
}

tcv_status tcv_degrid_tiled_half(tcv_ctx *ctx, tcv_plan *p, const float *d_grid, const void *d_Chalf,
                                 const float *d_planeScale, long long planeTaps, float *d_data)
{
    if (!ctx || !p || !d_grid || !d_Chalf || !d_planeScale || !d_data || planeTaps < 0 ||
        planeTaps > 0x7fffffffLL)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_degrid_tiled_half: bad argument");
    const HalfStorage h{1.f, d_planeScale, (int)planeTaps};
    return launch_degrid_tiled(ctx, p, (const float2 *)d_grid, d_Chalf, (float2 *)d_data, &h);
}

// ---- batching ------------------------------------------------------------------

tcv_status tcv_batch_create(tcv_ctx *ctx, int nJobs, const tcv_job *jobs, tcv_batch **out)
{
    if (!ctx || !out || nJobs < 1 || !jobs) return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_batch_create: bad argument");
    *out = nullptr;
    tcv_batch *b = new tcv_batch();
    b->nJobs = nJobs;
    for (int j = 0; j < nJobs; j++) {
        const tcv_job &J = jobs[j];
        if (!J.plan || !J.d_C || !J.d_grid) {
            delete b;
            return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_batch_create: job " + std::to_string(j) + " incomplete");
        }
        if (j == 0) {
            b->tw = J.plan->sh.tw;
            b->th = J.plan->sh.th;
            b->kern = J.plan->sh.kern;
            b->dkern = J.plan->sh.dkern;
            b->gridCols = J.plan->sh.gridCols;
        } else if (J.plan->sh.tw != b->tw || J.plan->sh.th != b->th || J.plan->sh.kern != b->kern ||
                   J.plan->sh.dkern != b->dkern || J.plan->sh.gridCols != b->gridCols) {
            delete b;
            return tcv_fail(ctx, TCV_ERR_INVALID,
                            "tcv_batch_create: all plans of a batch need one tile shape, kernel and grid_cols");
        }
        b->plans.push_back(J.plan);
        b->host.push_back(make_job_desc(J.plan, true, J.d_data, J.d_C, (float2 *)J.d_grid, (float2 *)J.d_out, nullptr));
        b->hostD.push_back(make_job_desc(J.plan, false, J.d_data, J.d_C, (float2 *)J.d_grid, (float2 *)J.d_out, nullptr));
    }
    // Jobs whose grids overlap (several channels into one grid, say) run in one
    // persistent launch: a work item that is the only one of its tile within
    // its own job may still meet another job's item on the same cells, so those
    // jobs flush every tile with atomics.
    mark_shared_grids(b->host);
    auto fail = [&](const char *m) {
        if (b->d_jobs) (void)hipFree(b->d_jobs);
        if (b->d_jobsD) (void)hipFree(b->d_jobsD);
        if (b->d_prefix) (void)hipFree(b->d_prefix);
        if (b->d_queue) (void)hipFree(b->d_queue);
        delete b;
        return tcv_fail(ctx, TCV_ERR_NOMEM, m);
    };
    if (hipMalloc((void **)&b->d_jobs, sizeof(JobDesc) * nJobs) != hipSuccess ||
        hipMalloc((void **)&b->d_jobsD, sizeof(JobDesc) * nJobs) != hipSuccess ||
        hipMalloc((void **)&b->d_prefix, sizeof(int) * 9 * (nJobs + 1)) != hipSuccess ||
        hipMalloc((void **)&b->d_queue, kBatchQueueBytes) != hipSuccess)
        return fail("tcv_batch_create: allocation failed");
    if (hipMemcpy(b->d_jobs, b->host.data(), sizeof(JobDesc) * nJobs, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_jobsD, b->hostD.data(), sizeof(JobDesc) * nJobs, hipMemcpyHostToDevice) != hipSuccess)
        return fail("tcv_batch_create: upload failed");
    *out = b;
    return TCV_OK;
}

tcv_status tcv_batch_destroy(tcv_ctx *ctx, tcv_batch *b)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!b) return TCV_OK;
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    (void)hipFree(b->d_jobs);
    (void)hipFree(b->d_jobsD);
    (void)hipFree(b->d_prefix);
    (void)hipFree(b->d_queue);
    delete b;
    return TCV_OK;
}

tcv_status tcv_batch_rebuild(tcv_ctx *ctx, tcv_batch *b)
{
    if (!ctx || !b) return TCV_ERR_INVALID;
    for (tcv_plan *p : b->plans) {
        if (!p->hasGeom) return tcv_fail(ctx, TCV_ERR_INVALID, "plan has no geometry");
        tcv_status st = plan_build_kernels(ctx, p, p->geom);
        if (st != TCV_OK) return st;
    }
    return TCV_OK;
}

tcv_status tcv_batch_grid(tcv_ctx *ctx, tcv_batch *b)
{
    if (!ctx || !b) return TCV_ERR_INVALID;
    for (const JobDesc &d : b->host)
        if (!d.data) return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_batch_grid: a job has no visibilities");
    return launch_batch(ctx, b, false);
}

tcv_status tcv_batch_degrid(tcv_ctx *ctx, tcv_batch *b)
{
    if (!ctx || !b) return TCV_ERR_INVALID;
    for (const JobDesc &d : b->hostD)
        if (!d.out) return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_batch_degrid: a job has no output");
    return launch_batch(ctx, b, true);
}

static tcv_status batch_enqueue(tcv_ctx *ctx, tcv_batch *b, int ops)
{
    tcv_status st = TCV_OK;
    if ((ops & 1) && (st = tcv_batch_rebuild(ctx, b)) != TCV_OK) return st;
    if ((ops & 2) && (st = tcv_batch_grid(ctx, b)) != TCV_OK) return st;
    if ((ops & 4) && (st = tcv_batch_degrid(ctx, b)) != TCV_OK) return st;
    return st;
}

tcv_status tcv_batch_run_graph(tcv_ctx *ctx, tcv_batch *b, int ops)
{
    if (!ctx || !b || ops < 1 || ops > 7) return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_batch_run_graph: bad argument");
    if (!b->graph || b->graphOps != ops) {
        if (b->graph) {
            (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
        const bool prof = ctx->profiling;
        ctx->profiling = false;   // no event records inside the capture
        TCV_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        tcv_status st = batch_enqueue(ctx, b, ops);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        ctx->profiling = prof;
        if (st != TCV_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return tcv_fail(ctx, TCV_ERR_DEVICE, std::string("capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->graph = nullptr;
            return tcv_fail(ctx, TCV_ERR_DEVICE, std::string("instantiate: ") + hipGetErrorString(e));
        }
        b->graphOps = ops;
    }
    ProfScope ps(ctx, "batch_graph");
    TCV_HIP(ctx, hipGraphLaunch(b->graph, ctx->stream));
    return TCV_OK;
}

// ---- C(w) on the device -----------------------------------------------------------

tcv_status tcv_attic_initC_device(tcv_ctx *ctx, double freq0, double cellSize, int wSize, int baseline,
                                  int overSample, int *support, double *wCellSize, float *d_C)
{
    if (!ctx || !support || !wCellSize || wSize < 1 || overSample < 1 || cellSize <= 0 || !d_C)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_attic_initC_device: bad argument");
    tcv_status st = tcv_attic_initC(freq0, cellSize, wSize, baseline, overSample, support, wCellSize, nullptr);
    if (st != TCV_OK) return st;
    double *d_sum = reinterpret_cast<double *>(ctx->d_counters + 4);
    return launch_attic_taps(ctx, freq0, cellSize, wSize, 2 * *support + 1, overSample, *wCellSize,
                             (float2 *)d_C, d_sum);
}

// ---- autotuner ---------------------------------------------------------------------

}  // extern "C"

std::string tuning_class(const tcv_geometry *g)
{
    int maxS = 0;
    for (int k = 0; k < g->nPlanes; k++) maxS = std::max(maxS, g->sSize[k]);
    int lv = 0;
    while ((1LL << (lv + 1)) <= std::max(1, g->nVis)) lv++;
    int lc = 0;
    while ((1LL << (lc + 1)) <= std::max<long long>(1, g->cSize)) lc++;
    return std::to_string(g->gSize) + "/" + std::to_string(maxS) + "/" + std::to_string(lc) + "/" +
           std::to_string(lv) + "/" + std::to_string(g->d_wPlane ? 1 : 0);
}

// Tile shapes worth trying for the widest kernel of a job.
static std::vector<std::pair<int, int>> candidate_tiles(int maxS)
{
    if (maxS >= 64) return {{128, 64}, {64, 128}, {32, 256}, {64, 64}, {64, 32}};
    if (maxS >= 48) return {{64, 64}, {64, 32}, {32, 64}};
    if (maxS >= 24) return {{32, 64}, {32, 32}, {64, 32}};
    if (maxS >= 12) return {{16, 64}, {16, 32}, {32, 32}};
    return {{8, 64}, {8, 32}, {16, 32}};
}

extern "C" {

tcv_status tcv_autotune(tcv_ctx *ctx, const tcv_geometry *g, const float *d_data, const float *d_C,
                        float *d_grid, float *d_out, int op, const tcv_tuning *candidates,
                        int nCandidates, int reps, tcv_tuning *best, double *bestMs)
{
    int maxS = 0;
    tcv_status st = check_geometry(ctx, g, &maxS);
    if (st != TCV_OK) return st;
    if (op < 1 || op > 3 || reps < 1 || !best || !d_C || ((op & 1) && (!d_data || !d_grid)) ||
        ((op & 2) && (!d_grid || !d_out)))
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_autotune: bad argument");
    double bestT = 1e300;
    bool any = false;
    // times one tuning on the job; false when it cannot be planned
    auto measure = [&](const tcv_tuning &t, double *msOut) -> tcv_status {
        *msOut = -1.0;
        tcv_plan *p = nullptr;
        if (tcv_plan_create(ctx, g, &t, &p) != TCV_OK) return TCV_OK;   // shape does not fit
        auto once = [&]() -> tcv_status {
            tcv_status s2 = TCV_OK;
            if (op & 1) s2 = tcv_grid_tiled(ctx, p, d_data, d_C, d_grid);
            if (s2 == TCV_OK && (op & 2)) s2 = tcv_degrid_tiled(ctx, p, d_grid, d_C, d_out);
            return s2;
        };
        tcv_status s3 = once();
        double ms = 0;
        if (s3 == TCV_OK) s3 = tcv_timer_start(ctx);
        for (int r = 0; r < reps && s3 == TCV_OK; r++) s3 = once();
        if (s3 == TCV_OK) s3 = tcv_timer_stop(ctx, &ms);
        tcv_plan_destroy(ctx, p);
        if (s3 == TCV_OK) *msOut = ms / reps;
        return s3;
    };
    auto consider = [&](const tcv_tuning &t) -> tcv_status {
        double ms;
        tcv_status s4 = measure(t, &ms);
        if (s4 == TCV_OK && ms >= 0 && ms < bestT) {
            bestT = ms;
            *best = t;
            any = true;
        }
        return s4;
    };
    if (candidates) {
        for (int k = 0; k < nCandidates; k++)
            if ((st = consider(candidates[k])) != TCV_OK) return st;
    } else {
        // staged search from the defaults: tile shape x kernel family, then the
        // pair gridder's column groups, the phase counts of the degridder and
        // gridder items, then the item size
        tcv_tuning d{0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0};
        auto_tuning(maxS, g, &d);
        int minS = maxS;
        for (int k = 0; k < g->nPlanes; k++) minS = std::min(minS, g->sSize[k]);
        for (auto &ts : candidate_tiles(maxS))
            for (auto kk : {std::make_pair(3, 3), std::make_pair(1, 2), std::make_pair(2, 3)}) {
                // tiles wider or taller than 64 have only the pair kernels
                if ((ts.first > 64 || ts.second > 64) && (kk.first != 3 || kk.second != 3 || minS < 2))
                    continue;
                tcv_tuning t = d;
                t.tile_w = ts.first;
                t.tile_h = ts.second;
                t.kernel = kk.first;
                t.degrid_kernel = kk.second;
                t.phases = 0;
                t.grid_phases = 0;
                t.xcd_split = -1;
                t.grid_cols = 0;
                auto_tuning(maxS, g, &t);   // phases and column groups for this tile shape
                if ((st = consider(t)) != TCV_OK) return st;
            }
        if (!any) return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tcv_autotune: no candidate could be planned");
        const tcv_tuning b0 = *best;
        if ((op & 1) && b0.kernel == 3)
            for (int c : {b0.tile_w, b0.tile_w / 2, b0.tile_w / 4}) {
                if (c == b0.grid_cols || !grid_cols_supported(b0.tile_w, b0.tile_h, c)) continue;
                tcv_tuning t = b0;
                t.grid_cols = c;
                if ((st = consider(t)) != TCV_OK) return st;
            }
        const tcv_tuning b1 = *best;
        // phase counts from 0.6x to 4x the defaults (multiples of 8 from 8 up), the
        // gridder's first, then the degridder's: the defaults follow the table's
        // size, and a job that uses a part of its table (w-planes of a narrower
        // range than the table spans) keeps its slices in the L2s with finer phases
        auto near = [](int v, double f) {
            const int x = std::max(1, (int)(v * f + 0.5));
            return x >= 8 ? x / 8 * 8 : x;
        };
        if (op & 1)
            for (double f : {0.6, 0.75, 1.5, 2.0, 3.0, 4.0}) {
                const int gp = near(b1.grid_phases, f);
                if (gp > b1.slice_keys || gp == b1.grid_phases) continue;
                tcv_tuning t = b1;
                t.grid_phases = gp;
                if ((st = consider(t)) != TCV_OK) return st;
            }
        const tcv_tuning bg = *best;
        if (op & 2)
            for (double f : {0.75, 1.25, 2.0, 4.0}) {
                const int p = near(bg.phases, f);
                if (p > bg.slice_keys || p == bg.phases) continue;
                tcv_tuning t = bg;
                t.phases = p;
                t.xcd_split = (p >= 8 && p % 8 == 0) ? 1 : 0;
                if ((st = consider(t)) != TCV_OK) return st;
            }
        const tcv_tuning b2 = *best;
        for (int ch : {512, 8192}) {
            tcv_tuning t = b2;
            t.chunk = ch;
            if ((st = consider(t)) != TCV_OK) return st;
        }
    }
    if (!any) return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tcv_autotune: no candidate could be planned");
    if (bestMs) *bestMs = bestT;
    ctx->tuned[tuning_class(g)] = *best;
    return TCV_OK;
}

}  // extern "C"
