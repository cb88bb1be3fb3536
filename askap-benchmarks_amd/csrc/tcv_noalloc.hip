// tcv_grid_noalloc: tcv_grid_plain's arguments and semantics, computed by the
// tiled path with its plan (entries, items, key counters, scan partials, queue
// heads) in a fixed device array instead of device memory allocated per call.
// For callers that cannot allocate device memory (code under stream capture,
// tight serving loops).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tcv_internal.h"

namespace {

// Every GPU the library runs on holds one copy of this array until the process
// ends.  Host code never uses its name as an address: the per-device address
// comes from hipGetSymbolAddress (noalloc_setup) and is kept in the context.
static __device__ __align__(256) unsigned char scratch[TCV_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];
hipError_t g_err[kMaxDevices];
bool g_done[kMaxDevices];

}  // namespace

// Called by tcv_create with the context's device current: the device address
// of `scratch` on that device, looked up once per device.  A failed lookup does
// not fail the context; tcv_grid_noalloc reports it.
void noalloc_setup(tcv_ctx *ctx)
{
    ctx->scratch = nullptr;
    ctx->scratchErr.clear();
    if (ctx->device < 0 || ctx->device >= kMaxDevices) {
        ctx->scratchErr = "device index beyond the scratch table";
        return;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    const int dev = ctx->device;
    if (!g_done[dev]) {
        void *p = nullptr;
        g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[dev] = g_err[dev] == hipSuccess ? p : nullptr;
        if (g_err[dev] != hipSuccess) (void)hipGetLastError();
        g_done[dev] = true;
    }
    ctx->scratch = g_ptr[dev];
    if (g_err[dev] != hipSuccess)
        ctx->scratchErr = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_err[dev]);
}

extern "C" tcv_status tcv_grid_noalloc(tcv_ctx *ctx, const tcv_geometry *g, const float *d_data,
                                       const float *d_C, float *d_grid)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!g) return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_grid_noalloc: null geometry");
    if (g->nVis == 0) return TCV_OK;
    if (!d_data || !d_C || !d_grid)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_grid_noalloc: null data, C or grid");
    if (!ctx->scratch)
        return tcv_fail(ctx, TCV_ERR_DEVICE, "tcv_grid_noalloc: no scratch address (" + ctx->scratchErr + ")");
    // the default plan, with C-slice keys halved until its buffers fit the
    // scratch array (the (tile, key) counters grow with the key count)
    tcv_plan p;
    size_t bytes = 0;
    for (int keys = 2048;; keys /= 2) {
        p = tcv_plan{};
        const tcv_tuning t{0, 0, 0, keys, 0, -1, 0, 0, 0, 0, 0, 0};
        tcv_status st = plan_setup(ctx, g, &t, &p, &bytes);
        if (st != TCV_OK) return st;
        if (bytes <= TCV_NOALLOC_SCRATCH_BYTES || keys == 1) break;
    }
    if (bytes > TCV_NOALLOC_SCRATCH_BYTES)
        return tcv_fail(ctx, TCV_ERR_UNSUPPORTED,
                        "tcv_grid_noalloc: this job's plan needs " + std::to_string(bytes) +
                            " bytes of scratch, more than the " +
                            std::to_string((size_t)TCV_NOALLOC_SCRATCH_BYTES) + " the library holds");
    plan_bind(&p, static_cast<char *>(ctx->scratch));
    p.ownsBlock = false;
    // queued behind every earlier call on this stream, which may still read the
    // previous call's plan from the same array
    if (g->d_wPlane)
        TCV_HIP(ctx, hipMemcpyAsync(p.d_sSize, g->sSize, sizeof(int) * g->nPlanes, hipMemcpyHostToDevice,
                                    ctx->stream));
    tcv_status st = tcv_plan_build(ctx, &p, g);   // clears the counters it accumulates into
    if (st != TCV_OK) return st;
    return launch_grid_tiled(ctx, &p, d_data, d_C, reinterpret_cast<float2 *>(d_grid), nullptr);
}
