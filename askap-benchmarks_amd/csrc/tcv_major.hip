// Major-cycle neighbours of the gridder on the device
// (current/tMajorACC/tMajorACC.cc): the shifted 2-D FFT (fftExec), fftFix,
// Hogbom CLEAN (findPeak / subtractPsf / deconvolve) and the visibility
// subtraction.  Everything stays resident in HBM; the Hogbom loop runs without
// a host round trip per iteration.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>

#include "tcv_internal.h"

namespace {

// ---- shifted FFT ----------------------------------------------------------------
// fftExec rotates the grid by g/2 on both axes, transforms, and rotates back
// (tMajorACC.cc:509-535).  For an even square grid that composition equals
// M . F . M, where M multiplies pixel (x, y) by (-1)^(x+y): rolling by g/2
// before a length-g DFT multiplies bin k by (-1)^k, and rolling after it is
// the DFT of the (-1)^n-modulated input times (-1)^(g/2); on a square grid the
// two (-1)^(g/2) factors cancel.  M is exact (a sign flip), so the shifts cost
// one in-place pass each instead of two copies through a second buffer.
// mode 0: x *= M; mode 1: x = M*x with fftFix folded in (real * scale, imag 0).
__global__ __launch_bounds__(256) void k_checker(float2 *__restrict__ g, int gSize, int mode,
                                                 float scale)
{
    const int y = blockIdx.y;
    const int x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 2;
    if (x0 >= gSize) return;
    float4 *p = reinterpret_cast<float4 *>(g + (size_t)y * gSize + x0);   // gSize even
    float4 v = *p;
    const float s = ((x0 + y) & 1) ? -1.f : 1.f;   // pixel x0; x0 + 1 has the other sign
    if (mode == 0) {
        v.x *= s;
        v.y *= s;
        v.z = -s * v.z;
        v.w = -s * v.w;
    } else {
        v.x = (s * v.x) * scale;
        v.y = 0.f;
        v.z = (-s * v.z) * scale;
        v.w = 0.f;
    }
    *p = v;
}

// fftFix on its own: grid = real(grid) * scale (tMajorACC.cc:620-627).
__global__ __launch_bounds__(256) void k_fft_fix(float2 *__restrict__ g, long long n, float scale)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        float2 v = g[i];
        g[i] = make_float2(v.x * scale, 0.f);
    }
}

// data -= model (tMajorACC.cc:1549-1556), complex, 16 B per thread per step.
__global__ __launch_bounds__(256) void k_vis_subtract(float2 *__restrict__ d,
                                                      const float2 *__restrict__ m, long long n)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const float2 a = d[i], b = m[i];
        d[i] = make_float2(a.x - b.x, a.y - b.y);
    }
}

// ---- Hogbom ------------------------------------------------------------------------
// findPeak picks the first pixel of largest |re| (tMajorACC.cc:689-701).  A
// 64-bit key orders pixels the same way under an unsigned max: |v|'s bits (31,
// non-negative floats order like their bit patterns), then the complemented
// index (smaller index wins a tie), then the sign of v (so the signed value
// can be rebuilt from the winning key).  One atomicMax per workgroup.
__device__ __forceinline__ unsigned long long peak_key(float v, unsigned idx)
{
    const unsigned a = __float_as_uint(v) & 0x7fffffffu;
    return ((unsigned long long)a << 33) | ((unsigned long long)(~idx) << 1) |
           (unsigned long long)(v < 0.f ? 1u : 0u);
}
__device__ __forceinline__ unsigned key_index(unsigned long long k)
{
    return ~(unsigned)((k >> 1) & 0xffffffffull);
}
__device__ __forceinline__ float key_value(unsigned long long k)
{
    const float a = __uint_as_float((unsigned)(k >> 33));
    return (k & 1ull) ? -a : a;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(k, o, 64);
        k = t > k ? t : k;
    }
    return k;
}

// workgroup max of k, then one atomicMax into *out (256 threads)
__device__ __forceinline__ void block_peak(unsigned long long k, unsigned long long *out)
{
    __shared__ unsigned long long s[4];
    k = wave_max_u64(k);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s[wave] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = s[0];
        for (int w = 1; w < 4; ++w) m = s[w] > m ? s[w] : m;
        if (m) atomicMax(out, m);
    }
}

// real part of a complex image into a float plane (and back)
__global__ __launch_bounds__(256) void k_real_plane(const float2 *__restrict__ src,
                                                    float *__restrict__ dst, long long n)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i].x;
}
__global__ __launch_bounds__(256) void k_real_writeback(float2 *__restrict__ dst,
                                                        const float *__restrict__ src, long long n)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i].x = src[i];
}

// peak of a float plane into *key (zeroed beforehand)
__global__ __launch_bounds__(256) void k_peak(const float *__restrict__ img, unsigned n,
                                              unsigned long long *key)
{
    unsigned long long best = 0;
    for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const unsigned long long k = peak_key(img[i], i);
        best = k > best ? k : best;
    }
    block_peak(best, key);
}

struct HogbomArgs {
    float *res;                  // residual real plane, W x W
    const float *psf;            // PSF real plane, PW x PW
    float2 *model;               // complex model image, W x W
    unsigned long long *keys;    // [0..2] rotating residual peaks, [3] PSF peak
    int W, PW;
    float gain;
};

// One Hogbom iteration (tMajorACC.cc:823-838) fused with the next findPeak:
// reads the residual peak of iteration i from keys[i % 3], adds gain * peak to
// the model, subtracts gain * peak * PSF over the overlap (subtractPsf,
// tMajorACC.cc:723-750, including its row offset diffy = ry - px), and
// reduces the updated residual's peak into keys[(i + 1) % 3].  keys[(i + 2) % 3]
// (read by iteration i - 1, written by iteration i + 1) is cleared here.
// Arithmetic is rounded exactly as the reference's float expressions.
template <bool VEC>
__global__ __launch_bounds__(256) void k_hogbom_step(HogbomArgs a, int i)
{
    // no fused multiply-add: each product and difference is rounded on its own,
    // as in the reference's float expressions (plain operators under this
    // pragma; helpers defined elsewhere would carry their own contract flags)
#pragma clang fp contract(off)
    const unsigned long long kc = a.keys[i % 3];
    const unsigned long long kp = a.keys[3];
    const float peak = key_value(kc);
    const int W = a.W, PW = a.PW;
    const unsigned pos = key_index(kc), ppos = key_index(kp);
    const int rx = (int)(pos % (unsigned)W), ry = (int)(pos / (unsigned)W);
    const int px = (int)(ppos % (unsigned)PW), py = (int)(ppos / (unsigned)PW);
    const int diffx = rx - px, diffy = ry - px;   // as in the reference
    const int startx = max(0, rx - px), starty = max(0, ry - py);
    const int stopx = min(W - 1, rx + (PW - px - 1)), stopy = min(W - 1, ry + (PW - py - 1));
    const float amp = a.gain * peak;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        a.model[pos].x = a.model[pos].x + peak * a.gain;
        a.keys[(i + 2) % 3] = 0ull;
    }
    // rows split over blockIdx.y, columns over blockIdx.x: no per-pixel division.
    // kRows rows a step, all their loads issued before any use, so that a
    // thread has that many reads in flight (one at a time left it waiting on
    // memory latency row after row).
    constexpr int kRows = 8;
    unsigned long long best = 0;
    if constexpr (VEC) {
        // W % 4 == 0: four adjacent pixels a thread, one 16-byte residual load
        // and store per row (the PSF row is read per pixel: any alignment)
        for (int x4 = (blockIdx.x * 256 + threadIdx.x) * 4; x4 < W; x4 += gridDim.x * 1024) {
            for (int y0 = blockIdx.y; y0 < W; y0 += gridDim.y * kRows) {
                float4 v[kRows];
                float pv[kRows][4];
                bool sub[kRows][4];
#pragma unroll
                for (int u = 0; u < kRows; ++u) {
                    const int y = y0 + u * (int)gridDim.y;
                    const int sy = y - diffy;
                    const bool rowOk = y < W;
                    const bool rowSub = rowOk && y >= starty && y <= stopy && sy >= 0 && sy < PW;
                    v[u] = rowOk ? *reinterpret_cast<const float4 *>(a.res + (size_t)y * W + x4)
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const int x = x4 + c;
                        sub[u][c] = rowSub && x >= startx && x <= stopx;
                        pv[u][c] = sub[u][c] ? a.psf[(ptrdiff_t)sy * PW - diffx + x] : 0.f;
                    }
                }
#pragma unroll
                for (int u = 0; u < kRows; ++u) {
                    const int y = y0 + u * (int)gridDim.y;
                    if (y >= W) continue;
                    float vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
                    bool any = false;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        if (sub[u][c]) {
                            vv[c] = vv[c] - amp * pv[u][c];
                            any = true;
                        }
                        const unsigned long long k =
                            peak_key(vv[c], (unsigned)y * (unsigned)W + (unsigned)(x4 + c));
                        best = k > best ? k : best;
                    }
                    if (any)
                        *reinterpret_cast<float4 *>(a.res + (size_t)y * W + x4) =
                            make_float4(vv[0], vv[1], vv[2], vv[3]);
                }
            }
        }
        block_peak(best, a.keys + (i + 1) % 3);
        return;
    }
    for (int x = blockIdx.x * 256 + threadIdx.x; x < W; x += gridDim.x * 256) {
        const bool colIn = x >= startx && x <= stopx;
        for (int y0 = blockIdx.y; y0 < W; y0 += gridDim.y * kRows) {
            float v[kRows], pv[kRows];
            bool sub[kRows];
#pragma unroll
            for (int u = 0; u < kRows; ++u) {
                const int y = y0 + u * (int)gridDim.y;
                const int sy = y - diffy;
                const bool rowOk = y < W;
                sub[u] = rowOk && colIn && y >= starty && y <= stopy && sy >= 0 && sy < PW;
                v[u] = rowOk ? a.res[(size_t)y * W + x] : 0.f;
                pv[u] = sub[u] ? a.psf[(ptrdiff_t)sy * PW - diffx + x] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < kRows; ++u) {
                const int y = y0 + u * (int)gridDim.y;
                if (y >= W) continue;
                float vv = v[u];
                if (sub[u]) {
                    vv = vv - amp * pv[u];
                    a.res[(size_t)y * W + x] = vv;
                }
                const unsigned long long k = peak_key(vv, (unsigned)y * (unsigned)W + (unsigned)x);
                best = k > best ? k : best;
            }
        }
    }
    block_peak(best, a.keys + (i + 1) % 3);
}

int stream_blocks(tcv_ctx *ctx, long long n, int perThread = 1)
{
    const long long want = (n + 256LL * perThread - 1) / (256LL * perThread);
    return (int)std::max<long long>(1, std::min<long long>(want, (long long)ctx->cus * 8));
}

}  // namespace

// ---- host side ---------------------------------------------------------------------

static tcv_status fft_plan(tcv_ctx *ctx, int gSize, hipfftHandle *out)
{
    auto it = ctx->fftPlans.find(gSize);
    if (it != ctx->fftPlans.end()) {
        *out = (hipfftHandle)it->second;
        return TCV_OK;
    }
    hipfftHandle h = nullptr;
    if (hipfftPlan2d(&h, gSize, gSize, HIPFFT_C2C) != HIPFFT_SUCCESS)
        return tcv_fail(ctx, TCV_ERR_DEVICE, "hipfftPlan2d failed for " + std::to_string(gSize));
    if (hipfftSetStream(h, ctx->stream) != HIPFFT_SUCCESS) {
        hipfftDestroy(h);
        return tcv_fail(ctx, TCV_ERR_DEVICE, "hipfftSetStream failed");
    }
    ctx->fftPlans[gSize] = (void *)h;
    *out = h;
    return TCV_OK;
}

void major_release(tcv_ctx *ctx)
{
    for (auto &kv : ctx->fftPlans) hipfftDestroy((hipfftHandle)kv.second);
    ctx->fftPlans.clear();
    if (ctx->majorWs) (void)hipFree(ctx->majorWs);
    ctx->majorWs = nullptr;
    ctx->majorWsBytes = 0;
}

// grow-only workspace: Hogbom's float planes and peak keys
static tcv_status major_ws(tcv_ctx *ctx, size_t bytes, void **p)
{
    if (ctx->majorWsBytes < bytes) {
        if (ctx->majorWs) TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->majorWs) (void)hipFree(ctx->majorWs);
        ctx->majorWs = nullptr;
        ctx->majorWsBytes = 0;
        if (hipMalloc(&ctx->majorWs, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return tcv_fail(ctx, TCV_ERR_NOMEM, "Hogbom workspace of " + std::to_string(bytes) +
                                                    " bytes");
        }
        ctx->majorWsBytes = bytes;
    }
    *p = ctx->majorWs;
    return TCV_OK;
}

static tcv_status fft_run(tcv_ctx *ctx, float *d_grid, int gSize, int forward, int fix, float scale)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!d_grid || gSize < 2 || (gSize & 1) || gSize > 65534)
        return tcv_fail(ctx, TCV_ERR_INVALID, "fftExec: gSize must be even (the shift is by gSize/2)");
    hipfftHandle h;
    tcv_status st = fft_plan(ctx, gSize, &h);
    if (st != TCV_OK) return st;
    ProfScope ps(ctx, "fft");
    float2 *g = reinterpret_cast<float2 *>(d_grid);
    const dim3 grid((gSize / 2 + 255) / 256, gSize);
    hipLaunchKernelGGL(k_checker, grid, dim3(256), 0, ctx->stream, g, gSize, 0, 1.f);
    TCV_HIP(ctx, hipGetLastError());
    if (hipfftExecC2C(h, (hipfftComplex *)g, (hipfftComplex *)g,
                      forward ? HIPFFT_FORWARD : HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
        return tcv_fail(ctx, TCV_ERR_DEVICE, "hipfftExecC2C failed");
    hipLaunchKernelGGL(k_checker, grid, dim3(256), 0, ctx->stream, g, gSize, fix ? 1 : 0, scale);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

extern "C" {

tcv_status tcv_fft_exec(tcv_ctx *ctx, float *d_grid, int gSize, int forward)
{
    return fft_run(ctx, d_grid, gSize, forward, 0, 1.f);
}

tcv_status tcv_fft_exec_fix(tcv_ctx *ctx, float *d_grid, int gSize, int forward, float scale)
{
    return fft_run(ctx, d_grid, gSize, forward, 1, scale);
}

tcv_status tcv_fft_fix(tcv_ctx *ctx, float *d_grid, long long nPixels, float scale)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (nPixels < 0 || (nPixels && !d_grid)) return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_fft_fix");
    if (!nPixels) return TCV_OK;
    ProfScope ps(ctx, "fft_fix");
    hipLaunchKernelGGL(k_fft_fix, dim3(stream_blocks(ctx, nPixels, 4)), dim3(256), 0, ctx->stream,
                       (float2 *)d_grid, nPixels, scale);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status tcv_vis_subtract(tcv_ctx *ctx, float *d_data, const float *d_model, long long nVis)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (nVis < 0 || (nVis && (!d_data || !d_model)))
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_vis_subtract");
    if (!nVis) return TCV_OK;
    ProfScope ps(ctx, "vis_subtract");
    hipLaunchKernelGGL(k_vis_subtract, dim3(stream_blocks(ctx, nVis, 4)), dim3(256), 0,
                       ctx->stream, (float2 *)d_data, (const float2 *)d_model, nVis);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status tcv_find_peak(tcv_ctx *ctx, const float *d_image, long long nPixels, float *maxVal,
                         long long *maxPos)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!d_image || nPixels < 1 || nPixels > 0xffffffffLL || !maxVal || !maxPos)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_find_peak: bad argument (1 <= nPixels < 2^32)");
    const size_t planeB = ((size_t)nPixels * 4 + 255) & ~size_t(255);
    void *ws;
    tcv_status st = major_ws(ctx, planeB + 256, &ws);
    if (st != TCV_OK) return st;
    float *plane = (float *)ws;
    unsigned long long *key = (unsigned long long *)((char *)ws + planeB);
    ProfScope ps(ctx, "find_peak");
    st = launch_fill_u32(ctx, key, 0u, 2);
    if (st != TCV_OK) return st;
    hipLaunchKernelGGL(k_real_plane, dim3(stream_blocks(ctx, nPixels, 4)), dim3(256), 0,
                       ctx->stream, (const float2 *)d_image, plane, nPixels);
    hipLaunchKernelGGL(k_peak, dim3(stream_blocks(ctx, nPixels, 4)), dim3(256), 0, ctx->stream,
                       plane, (unsigned)nPixels, key);
    TCV_HIP(ctx, hipGetLastError());
    unsigned long long k = 0;
    TCV_HIP(ctx, hipMemcpyAsync(&k, key, 8, hipMemcpyDeviceToHost, ctx->stream));
    TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const unsigned a = (unsigned)(k >> 33);
    float av;
    std::memcpy(&av, &a, 4);
    *maxVal = (k & 1) ? -av : av;
    *maxPos = k ? (long long)(~(unsigned)((k >> 1) & 0xffffffffull)) : 0;
    return TCV_OK;
}

tcv_status tcv_hogbom(tcv_ctx *ctx, float *d_residual, long long width, const float *d_psf,
                      long long psfWidth, float *d_model, int niters, float gain,
                      float *psfPeakVal, long long *psfPeakPos, float *dirtyPeakVal,
                      long long *dirtyPeakPos)
{
    if (!ctx) return TCV_ERR_INVALID;
    if (!d_residual || !d_psf || !d_model || width < 1 || psfWidth < 1 || niters < 0 ||
        width * width > 0xffffffffLL || psfWidth * psfWidth > 0xffffffffLL)
        return tcv_fail(ctx, TCV_ERR_INVALID, "tcv_hogbom: bad argument (widths below 65536)");
    const long long n = width * width, pn = psfWidth * psfWidth;
    const size_t resB = ((size_t)n * 4 + 255) & ~size_t(255);
    const size_t psfB = ((size_t)pn * 4 + 255) & ~size_t(255);
    void *ws;
    tcv_status st = major_ws(ctx, resB + psfB + 256, &ws);
    if (st != TCV_OK) return st;
    float *res = (float *)ws;
    float *psf = (float *)((char *)ws + resB);
    unsigned long long *keys = (unsigned long long *)((char *)ws + resB + psfB);
    ProfScope ps(ctx, "hogbom");
    st = launch_fill_u32(ctx, keys, 0u, 8);
    if (st != TCV_OK) return st;
    hipLaunchKernelGGL(k_real_plane, dim3(stream_blocks(ctx, n, 4)), dim3(256), 0, ctx->stream,
                       (const float2 *)d_residual, res, n);
    hipLaunchKernelGGL(k_real_plane, dim3(stream_blocks(ctx, pn, 4)), dim3(256), 0, ctx->stream,
                       (const float2 *)d_psf, psf, pn);
    hipLaunchKernelGGL(k_peak, dim3(stream_blocks(ctx, pn, 4)), dim3(256), 0, ctx->stream, psf,
                       (unsigned)pn, keys + 3);
    hipLaunchKernelGGL(k_peak, dim3(stream_blocks(ctx, n, 4)), dim3(256), 0, ctx->stream, res,
                       (unsigned)n, keys + 0);
    TCV_HIP(ctx, hipGetLastError());
    unsigned long long first = 0;
    if (dirtyPeakVal || dirtyPeakPos)
        TCV_HIP(ctx, hipMemcpyAsync(&first, keys, 8, hipMemcpyDeviceToHost, ctx->stream));
    HogbomArgs a{res, psf, (float2 *)d_model, keys, (int)width, (int)psfWidth, gain};
    // four pixels a thread when the rows are whole float4s
    const bool vec = width % 4 == 0;
    const long long perRow = vec ? width / 4 : width;
    const int bx = (int)std::min<long long>((perRow + 255) / 256, 64);
    const int by = (int)std::max<long long>(1, std::min<long long>(width, (long long)ctx->cus * 8 / bx));
    for (int i = 0; i < niters; ++i) {
        if (vec)
            hipLaunchKernelGGL(k_hogbom_step<true>, dim3(bx, by), dim3(256), 0, ctx->stream, a, i);
        else
            hipLaunchKernelGGL(k_hogbom_step<false>, dim3(bx, by), dim3(256), 0, ctx->stream, a, i);
    }
    TCV_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_real_writeback, dim3(stream_blocks(ctx, n, 4)), dim3(256), 0, ctx->stream,
                       (float2 *)d_residual, (const float *)res, n);
    TCV_HIP(ctx, hipGetLastError());
    unsigned long long pk = 0;
    if (psfPeakVal || psfPeakPos)
        TCV_HIP(ctx, hipMemcpyAsync(&pk, keys + 3, 8, hipMemcpyDeviceToHost, ctx->stream));
    if (psfPeakVal || psfPeakPos || dirtyPeakVal || dirtyPeakPos) {
        TCV_HIP(ctx, hipStreamSynchronize(ctx->stream));
        auto val = [](unsigned long long k) {
            const unsigned a2 = (unsigned)(k >> 33);
            float f;
            std::memcpy(&f, &a2, 4);
            return (k & 1) ? -f : f;
        };
        auto pos = [](unsigned long long k) {
            return k ? (long long)(~(unsigned)((k >> 1) & 0xffffffffull)) : 0LL;
        };
        if (psfPeakVal) *psfPeakVal = val(pk);
        if (psfPeakPos) *psfPeakPos = pos(pk);
        if (dirtyPeakVal) *dirtyPeakVal = val(first);
        if (dirtyPeakPos) *dirtyPeakPos = pos(first);
    }
    return TCV_OK;
}

}  // extern "C"
