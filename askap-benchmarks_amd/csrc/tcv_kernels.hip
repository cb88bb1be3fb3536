// HIP kernels for gfx950 (MI355X): FP64 index set-up, plain and tiled
// gridding/degridding, and the device-side binning plan.
//
// Design notes (the reference loops these replace are cited per kernel):
//  * wave64 everywhere; no MFMA -- gridding is a scatter-add, degridding a
//    gather-reduce, both bound by the C(w) stream (8 B per grid point).
//  * C(w) is read through raw buffer loads with a 32-bit byte offset; lanes
//    with nothing to do get an offset past num_records, which the buffer unit
//    returns as zero without touching memory, so masked lanes need no exec
//    juggling and can never read out of bounds.
//  * The tiled kernels keep the grid tile in VGPRs (owner computes) and
//    stream only C; the plain kernels are the one-atomic-per-point baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "tcv_internal.h"

namespace {

constexpr unsigned kOOB = 0x80000000u;  // byte offset beyond any C buffer we accept

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned bytes)
{
    // word3 0x00020000: raw 32-bit buffer format used for untyped loads on gfx9.
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, bytes, 0x00020000);
}

// Load of wave-uniform data that no wave writes while the kernel runs (work
// items, entries, visibilities, plane scales).  Through the constant address
// space the compiler issues a scalar (SMEM) load into SGPRs, instead of a
// vector load, a wait for it and readfirstlane.
template <typename T>
__device__ __forceinline__ T cload(const T *p, int i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T *)p)[i];
#else
    return p[i];   // the host pass only parses device functions
#endif
}

// a * b for 0 <= a, b < 2^24, b wave-uniform: the full-rate 24-bit multiply,
// written as the instruction so that the compiler cannot fold it into a
// neighbouring expression as a quarter-rate 32- or 64-bit multiply
__device__ __forceinline__ int umul24(int a, int b)
{
    int r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(b), "v"(a));
    return r;
}

// a * b + c (24-bit unsigned a, b; b wave-uniform)
__device__ __forceinline__ unsigned umad24(int a, int b, unsigned c)
{
    unsigned r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// a * b + c (24-bit signed a, b; b wave-uniform)
__device__ __forceinline__ unsigned imad24(int a, int b, unsigned c)
{
    unsigned r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

__device__ __forceinline__ float2 load_c(__amdgpu_buffer_rsrc_t r, unsigned voff)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}


// Sum over the 64 lanes of a wave (all lanes must be active).  Four DPP
// steps reduce each 16-lane row; four readlanes combine the rows.  The result
// is wave-uniform.
__device__ __forceinline__ float wave_sum(float x)
{
    auto dpp = [](float v, auto ctrl) {
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
            0, __builtin_bit_cast(int, v), decltype(ctrl)::value, 0xF, 0xF, false));
    };
    x += dpp(x, std::integral_constant<int, 0xB1>{});   // quad_perm [1,0,3,2]
    x += dpp(x, std::integral_constant<int, 0x4E>{});   // quad_perm [2,3,0,1]
    x += dpp(x, std::integral_constant<int, 0x141>{});  // row_half_mirror
    x += dpp(x, std::integral_constant<int, 0x140>{});  // row_mirror
    const int xi = __builtin_bit_cast(int, x);
    const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(xi, 0));
    const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(xi, 16));
    const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(xi, 32));
    const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(xi, 48));
    return (r0 + r1) + (r2 + r3);
}

// ---- fill (graph-safe zeroing) ------------------------------------------------

__global__ void k_fill_u32(unsigned *p, unsigned v, size_t n)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Streaming device copy, 16 B per lane and iteration: a known-size HBM read and
// write (the calibration kernel for hardware traffic counters).
__global__ __launch_bounds__(256) void k_copy_u4(uint4 *__restrict__ dst,
                                                  const uint4 *__restrict__ src, size_t n16)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

// Streaming read of a power-of-two buffer, 16 B per lane (the read-bandwidth
// probe): each round the grid reads one window of its own size (thread g reads
// element base + g, coalesced), and the window moves on by its size plus an odd
// offset, so no CU meets data its L1 still holds.  A buffer that fits the L2s
// measures their bandwidth; a large one, HBM's.  The sum only keeps the loads.
__global__ __launch_bounds__(256) void k_read_stream(const uint4 *__restrict__ p, unsigned mask,
                                                     unsigned adv, int rounds, unsigned *sink)
{
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned base = 0, acc = 0;
    for (int r = 0; r < rounds; r += 8) {
        uint4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = p[(base + g) & mask];
            base += adv;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    if (acc == 0x9e3779b9u) sink[0] = acc;
}

// ---- a6: FP64 index set-up ----------------------------------------------------
// attic/tConvolveMPI/Benchmark.cc:293-322 and current/tConvolveMPI/Benchmark.cc:610-643.
// Same expression order as the reference; contraction off so every product and
// quotient is rounded exactly as on the host.
__global__ __launch_bounds__(256) void k_offsets(
    const double *__restrict__ u, const double *__restrict__ v,
    const double *__restrict__ w, int nSamples, const double *__restrict__ freq,
    int nChan, double cellSize, double wCellSize, int wSize, int gSize,
    int overSample, const int *__restrict__ sSize, const int *__restrict__ cOffset0,
    int *__restrict__ iuOut, int *__restrict__ ivOut, int *__restrict__ wPlaneOut,
    int *__restrict__ cOffsetOut, unsigned long long *__restrict__ numPerPlane,
    unsigned long long *__restrict__ bad, int useLdsHist)
{
#pragma clang fp contract(off)
    extern __shared__ unsigned hist[];
    if (useLdsHist) {
        for (int k = threadIdx.x; k < wSize; k += blockDim.x) hist[k] = 0;
        __syncthreads();
    }
    const int total = nSamples * nChan;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int isBad = 0;
    if (idx < total) {
        const int i = idx / nChan;
        const int chan = idx - i * nChan;
        const double f = freq[chan];

        const double uScaled = f * u[i] / cellSize;
        int a = (int)uScaled;
        if (uScaled < (double)a) a -= 1;
        const int fracu = (int)(overSample * (uScaled - (double)a));
        const int iu = a + gSize / 2;

        const double vScaled = f * v[i] / cellSize;
        int b = (int)vScaled;
        if (vScaled < (double)b) b -= 1;
        const int fracv = (int)(overSample * (vScaled - (double)b));
        const int iv = b + gSize / 2;

        int woff;
        int s, co;
        if (cOffset0 == nullptr) {
            // attic: uniform width, planes laid out contiguously
            const double wScaled = f * w[i] / wCellSize;
            woff = wSize / 2 + (int)wScaled;
            s = sSize[0];
            co = s * s * (fracu + overSample * (fracv + overSample * woff));
        } else {
            woff = 0;
            if (wCellSize > 0.0) {
                const double wScaled = f * w[i] / wCellSize;
                woff = wSize / 2 + (int)wScaled;
            }
            const bool planeOk = woff >= 0 && woff < wSize;
            s = planeOk ? sSize[woff] : 0;
            co = planeOk ? s * s * (fracu + overSample * fracv) + cOffset0[woff] : 0;
        }
        const bool planeOk = woff >= 0 && woff < wSize;
        const int half = s / 2;
        // footprint rows [iv, iv + s), columns [iu - s/2, iu - s/2 + s)
        isBad = !planeOk || iu - half < 0 || iv < 0 || iu - half + s > gSize ||
                iv + s > gSize;
        iuOut[idx] = iu;
        ivOut[idx] = iv;
        cOffsetOut[idx] = co;
        if (wPlaneOut) wPlaneOut[idx] = woff;
        if (numPerPlane && planeOk) {
            if (useLdsHist) atomicAdd(&hist[woff], 1u);
            else atomicAdd(&numPerPlane[woff], 1ull);
        }
    }
    // one atomic per wave for the out-of-range count
    const unsigned long long nb = __builtin_popcountll(__ballot(isBad));
    if ((threadIdx.x & 63) == 0 && nb) atomicAdd(bad, nb);
    if (useLdsHist) {
        __syncthreads();
        for (int k = threadIdx.x; k < wSize; k += blockDim.x)
            if (hist[k]) atomicAdd(&numPerPlane[k], (unsigned long long)hist[k]);
    }
}

// Validates one visibility; returns its width and footprint corner.
struct VisGeom {
    int s, u0, v0, c0;
    bool ok;
};

__device__ __forceinline__ VisGeom vis_geom(const GeomDev &g, int d)
{
    VisGeom r;
    int s = g.sUniform;
    bool ok = true;
    if (g.wPlane) {
        const int p = g.wPlane[d];
        ok = p >= 0 && p < g.nPlanes;
        s = ok ? g.sSize[p] : 0;
    }
    // footprint rows [iv, iv + s), columns [iu - s/2, iu - s/2 + s): the
    // reference offsets only the column (gind = iu + gSize*iv - support,
    // attic/tConvolveMPI/Benchmark.cc:133)
    const int half = s / 2;
    r.s = s;
    r.u0 = g.iu[d] - half;
    r.v0 = g.iv[d];
    r.c0 = g.cOffset[d];
    ok = ok && s > 0 && r.u0 >= 0 && r.v0 >= 0 && r.u0 + s <= g.gSize &&
         r.v0 + s <= g.gSize && r.c0 >= 0 &&
         (long long)r.c0 + (long long)s * s <= g.cSize;
    r.ok = ok;
    return r;
}

// ---- a7 plain gridder ---------------------------------------------------------
// attic/tConvolveMPI/Benchmark.cc:131-154: one workgroup per visibility, its
// threads walk the flattened sSize x sSize footprint; every tap is one pair of
// FP32 global atomics (the one-job-per-launch baseline).
template <int BS>
__global__ __launch_bounds__(BS) void k_grid_plain(GeomDev g, const float2 *__restrict__ data,
                                                   const float2 *__restrict__ C,
                                                   float2 *__restrict__ grid,
                                                   unsigned long long *bad)
{
    const int d = blockIdx.x;
    const VisGeom vg = vis_geom(g, d);
    if (!vg.ok) {
        if (threadIdx.x == 0) atomicAdd(bad, 1ull);
        return;
    }
    const int s = vg.s;
    const float2 dv = data[d];
    const int n = s * s;
    int r = threadIdx.x / s, c = threadIdx.x - (threadIdx.x / s) * s;
    const int sr = BS / s, sc = BS - (BS / s) * s;
    const float2 *cp = C + vg.c0;
    for (int p = threadIdx.x; p < n; p += BS) {
        const float2 cv = cp[p];
        float *gp = reinterpret_cast<float *>(grid + ((size_t)(vg.v0 + r) * g.gSize + vg.u0 + c));
        unsafeAtomicAdd(gp, dv.x * cv.x - dv.y * cv.y);
        unsafeAtomicAdd(gp + 1, dv.x * cv.y + dv.y * cv.x);
        c += sc;
        r += sr;
        if (c >= s) { c -= s; r++; }
    }
}

// ---- a8 plain degridder -------------------------------------------------------
// attic/tConvolveMPI/Benchmark.cc:165-193: one wave64 per visibility, lanes walk
// the flattened footprint, then a DPP reduction.
__global__ __launch_bounds__(256) void k_degrid_plain(GeomDev g, const float2 *__restrict__ grid,
                                                      const float2 *__restrict__ C,
                                                      float2 *__restrict__ out,
                                                      unsigned long long *bad)
{
    const int lane = threadIdx.x & 63;
    const int d = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (d >= g.nVis) return;   // whole wave leaves together
    const VisGeom vg = vis_geom(g, d);
    if (!vg.ok) {
        if (lane == 0) {
            atomicAdd(bad, 1ull);
            out[d] = make_float2(0.f, 0.f);
        }
        return;
    }
    const int s = vg.s;
    const int n = s * s;
    int r = lane / s, c = lane - (lane / s) * s;
    const int sr = 64 / s, sc = 64 - (64 / s) * s;
    const float2 *cp = C + vg.c0;
    float re = 0.f, im = 0.f;
    for (int p = lane; p < n; p += 64) {
        const float2 cv = cp[p];
        const float2 gv = grid[(size_t)(vg.v0 + r) * g.gSize + vg.u0 + c];
        re += gv.x * cv.x - gv.y * cv.y;
        im += gv.x * cv.y + gv.y * cv.x;
        c += sc;
        r += sr;
        if (c >= s) { c -= s; r++; }
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) out[d] = make_float2(re, im);
}

// ---- FP64 reference mode of the plain kernels -----------------------------------
// The same footprint walks as k_grid_plain / k_degrid_plain on the same stored
// complex64 taps and visibilities, promoted to double: each tap product is
// exact in FP64 (24-bit x 24-bit significands) and sums accumulate in FP64
// (grid: FP64 atomics into a complex128 grid; degrid: per-lane FP64 sums and a
// wave reduction).  Every FP32 or FP16 result can then state its error against
// an FP64 gridding on the device, at any size.
__global__ __launch_bounds__(256) void k_grid_plain_f64(GeomDev g, const float2 *__restrict__ data,
                                                        const float2 *__restrict__ C,
                                                        double2 *__restrict__ grid,
                                                        unsigned long long *bad)
{
    constexpr int BS = 256;
    const int d = blockIdx.x;
    const VisGeom vg = vis_geom(g, d);
    if (!vg.ok) {
        if (threadIdx.x == 0) atomicAdd(bad, 1ull);
        return;
    }
    const int s = vg.s;
    const float2 dv = data ? data[d] : make_float2(1.f, 0.f);
    const double dr = dv.x, di = dv.y;
    const int n = s * s;
    int r = threadIdx.x / s, c = threadIdx.x - (threadIdx.x / s) * s;
    const int sr = BS / s, sc = BS - (BS / s) * s;
    const float2 *cp = C + vg.c0;
    for (int p = threadIdx.x; p < n; p += BS) {
        const float2 cv = cp[p];
        const double cr = cv.x, ci = cv.y;
        double *gp = reinterpret_cast<double *>(grid + ((size_t)(vg.v0 + r) * g.gSize + vg.u0 + c));
        unsafeAtomicAdd(gp, dr * cr - di * ci);
        unsafeAtomicAdd(gp + 1, dr * ci + di * cr);
        c += sc;
        r += sr;
        if (c >= s) { c -= s; r++; }
    }
}

__device__ __forceinline__ double wave_sum_f64(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__global__ __launch_bounds__(256) void k_degrid_plain_f64(GeomDev g, const float2 *__restrict__ grid,
                                                          const float2 *__restrict__ C,
                                                          double2 *__restrict__ out,
                                                          unsigned long long *bad)
{
    const int lane = threadIdx.x & 63;
    const int d = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (d >= g.nVis) return;   // whole wave leaves together
    const VisGeom vg = vis_geom(g, d);
    if (!vg.ok) {
        if (lane == 0) {
            atomicAdd(bad, 1ull);
            out[d] = make_double2(0.0, 0.0);
        }
        return;
    }
    const int s = vg.s;
    const int n = s * s;
    int r = lane / s, c = lane - (lane / s) * s;
    const int sr = 64 / s, sc = 64 - (64 / s) * s;
    const float2 *cp = C + vg.c0;
    double re = 0.0, im = 0.0;
    for (int p = lane; p < n; p += 64) {
        const float2 cv = cp[p];
        const float2 gv = grid[(size_t)(vg.v0 + r) * g.gSize + vg.u0 + c];
        re += (double)gv.x * cv.x - (double)gv.y * cv.y;
        im += (double)gv.x * cv.y + (double)gv.y * cv.x;
        c += sc;
        r += sr;
        if (c >= s) { c -= s; r++; }
    }
    re = wave_sum_f64(re);
    im = wave_sum_f64(im);
    if (lane == 0) out[d] = make_double2(re, im);
}

// ---- scan (exclusive, int32, total written at out[n]) -------------------------

constexpr int SCAN_BS = 256;
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IPT;

__device__ __forceinline__ int block_exclusive_scan(int x, int *total)
{
    __shared__ int wsum[SCAN_BS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < SCAN_BS / 64; ++k) {
        const int t = wsum[k];
        if (k < wv) before += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + incl - x;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_reduce(const int *__restrict__ in, long long n,
                                                         int *__restrict__ sums)
{
    const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_IPT;
    int t = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k)
        if (base + k < n) t += in[base + k];
    int total;
    (void)block_exclusive_scan(t, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_apply(const int *__restrict__ in, long long n,
                                                        const int *__restrict__ offs,
                                                        int *__restrict__ out)
{
    const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_IPT;
    int vals[SCAN_IPT];
    int t = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        vals[k] = (base + k < n) ? in[base + k] : 0;
        t += vals[k];
    }
    int total;
    int run = block_exclusive_scan(t, &total) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        if (base + k < n) out[base + k] = run;
        run += vals[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BS - 1) out[n] = run;
}

// ---- tiled path: plan kernels -------------------------------------------------

struct PlanDevArgs {
    int tw, th, nTX, k2, phases, chunk, nTiles, xcd;
    long long cSize;
};

// The cOffset range of the binned entries, for phases over the part of the
// table in use: meta[4] = max(0x7fffffff - c0), meta[5] = max(c0 + 1), so the
// zeroed counters read as an empty range.  One atomic pair per wave.
__device__ __forceinline__ void note_c_range(int cLo, int cHi, int *meta)
{
    for (int o = 32; o > 0; o >>= 1) {
        cLo = min(cLo, __shfl_xor(cLo, o, 64));
        cHi = max(cHi, __shfl_xor(cHi, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && cHi >= 0) {
        atomicMax(reinterpret_cast<unsigned *>(meta + 4), 0x7fffffffu - (unsigned)cLo);
        atomicMax(reinterpret_cast<unsigned *>(meta + 5), (unsigned)cHi + 1u);
    }
}

// pass 0 counts (tile, slice-key) buckets; pass 1 scatters the entries.
// A visibility with footprint [u0, u0+s) x [v0, v0+s) is entered once in every
// tile its footprint overlaps.  Entry = {u0 | v0 << 16, cOffset, vis, s}.
__global__ __launch_bounds__(256) void k_plan_bin(int pass, GeomDev g, PlanDevArgs a,
                                                  int *__restrict__ counts,
                                                  int *__restrict__ cursor,
                                                  int4 *__restrict__ entries,
                                                  int *__restrict__ meta)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    int isBad = 0;
    int cLo = 0x7fffffff, cHi = -1;
    if (d < g.nVis) {
        const VisGeom vg = vis_geom(g, d);
        if (!vg.ok) {
            isBad = 1;
        } else {
            cLo = cHi = vg.c0;
            const int key2 = (int)(((long long)vg.c0 * a.k2) / a.cSize);
            const int tx0 = vg.u0 / a.tw, tx1 = (vg.u0 + vg.s - 1) / a.tw;
            const int ty0 = vg.v0 / a.th, ty1 = (vg.v0 + vg.s - 1) / a.th;
            const int plane = g.wPlane ? g.wPlane[d] : 0;
            const int4 ent = make_int4(vg.u0 | (vg.v0 << 16), vg.c0, d, vg.s | (plane << 16));
            for (int ty = ty0; ty <= ty1; ++ty)
                for (int tx = tx0; tx <= tx1; ++tx) {
                    const long long key = (long long)(ty * a.nTX + tx) * a.k2 + key2;
                    if (pass == 0) {
                        atomicAdd(&counts[key], 1);
                    } else {
                        const int pos = atomicAdd(&cursor[key], 1);
                        entries[pos] = ent;
                    }
                }
        }
    }
    if (pass == 0) {
        const int nb = __builtin_popcountll(__ballot(isBad));
        if ((threadIdx.x & 63) == 0 && nb) atomicAdd(&meta[3], nb);
        note_c_range(cLo, cHi, meta);
    }
}

// The same binning with the (tile, key) histogram aggregated in LDS, for key
// spaces that fit (nKeys <= kLdsKeys): each block takes kBinVPT visibilities
// per thread, counts locally, and touches each global counter at most once
// (pass 0: add the block's count; pass 1: reserve the block's range, then place
// entries at range start + LDS rank).  Cuts the global atomics on hot tile
// counters by the number of visibilities per block.
constexpr int kBinVPT = 16;
constexpr int kLdsKeys = 6144;   // 2 x 6144 ints = 48 KB of LDS in pass 1

__global__ __launch_bounds__(256) void k_plan_bin_lds(int pass, GeomDev g, PlanDevArgs a,
                                                      int *__restrict__ counts,
                                                      int *__restrict__ cursor,
                                                      int4 *__restrict__ entries,
                                                      int *__restrict__ meta)
{
    extern __shared__ int lds[];
    const int nKeys = a.nTiles * a.k2;
    int *hist = lds;
    int *rank = lds + nKeys;
    for (int k = threadIdx.x; k < (pass ? 2 : 1) * nKeys; k += blockDim.x) lds[k] = 0;
    __syncthreads();
    const int base = blockIdx.x * blockDim.x * kBinVPT;
    int isBad = 0;
    int cLo = 0x7fffffff, cHi = -1;
    for (int r = 0; r < kBinVPT; ++r) {
        const int d = base + r * blockDim.x + threadIdx.x;
        if (d >= g.nVis) break;
        const VisGeom vg = vis_geom(g, d);
        if (!vg.ok) {
            isBad++;
            continue;
        }
        cLo = min(cLo, vg.c0);
        cHi = max(cHi, vg.c0);
        const int key2 = (int)(((long long)vg.c0 * a.k2) / a.cSize);
        const int tx0 = vg.u0 / a.tw, tx1 = (vg.u0 + vg.s - 1) / a.tw;
        const int ty0 = vg.v0 / a.th, ty1 = (vg.v0 + vg.s - 1) / a.th;
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) atomicAdd(&hist[(ty * a.nTX + tx) * a.k2 + key2], 1);
    }
    __syncthreads();
    if (pass == 0) {
        for (int k = threadIdx.x; k < nKeys; k += blockDim.x)
            if (hist[k]) atomicAdd(&counts[k], hist[k]);
        // reduce the out-of-range count over the wave, one atomic per wave
        int nb = isBad;
        for (int o = 32; o > 0; o >>= 1) nb += __shfl_xor(nb, o, 64);
        if ((threadIdx.x & 63) == 0 && nb) atomicAdd(&meta[3], nb);
        note_c_range(cLo, cHi, meta);
        return;
    }
    for (int k = threadIdx.x; k < nKeys; k += blockDim.x)
        if (hist[k]) hist[k] = atomicAdd(&cursor[k], hist[k]);
    __syncthreads();
    for (int r = 0; r < kBinVPT; ++r) {
        const int d = base + r * blockDim.x + threadIdx.x;
        if (d >= g.nVis) break;
        const VisGeom vg = vis_geom(g, d);
        if (!vg.ok) continue;
        const int key2 = (int)(((long long)vg.c0 * a.k2) / a.cSize);
        const int tx0 = vg.u0 / a.tw, tx1 = (vg.u0 + vg.s - 1) / a.tw;
        const int ty0 = vg.v0 / a.th, ty1 = (vg.v0 + vg.s - 1) / a.th;
        const int plane = g.wPlane ? g.wPlane[d] : 0;
        const int4 ent = make_int4(vg.u0 | (vg.v0 << 16), vg.c0, d, vg.s | (plane << 16));
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) {
                const int key = (ty * a.nTX + tx) * a.k2 + key2;
                entries[hist[key] + atomicAdd(&rank[key], 1)] = ent;
            }
    }
}

// Second level of the two-level binning: one workgroup per tile orders the
// tile's entries (grouped by tile by the first level, in `tmp`) by C-slice
// key with an LDS counting sort, writes them to `entries` and the tile's key
// offsets to starts[tile * k2 + key].  Block 0 also writes the total at
// starts[nTiles * k2].
__global__ __launch_bounds__(256) void k_plan_tile_keys(PlanDevArgs a, const int *__restrict__ tileStart,
                                                        const int4 *__restrict__ tmp,
                                                        int4 *__restrict__ entries,
                                                        int *__restrict__ starts)
{
    extern __shared__ int lds[];
    int *hist = lds, *cur = lds + a.k2;
    const int t = blockIdx.x;
    const int b = tileStart[t], e = tileStart[t + 1];
    // key = floor(cOffset * k2 / cSize) up to rounding: monotone in cOffset and
    // computed the same way in both passes (an FP64 product, no division)
    const double scale = (double)a.k2 / (double)a.cSize;
    auto key_of = [&](int c0) { return min((int)((double)c0 * scale), a.k2 - 1); };
    for (int k = threadIdx.x; k < a.k2; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) atomicAdd(&hist[key_of(tmp[i].y)], 1);
    __syncthreads();
    // exclusive scan of hist: each thread a run of consecutive keys
    const int per = (a.k2 + blockDim.x - 1) / blockDim.x;
    const int k0 = min((int)threadIdx.x * per, a.k2), k1 = min(k0 + per, a.k2);
    int run = 0;
    for (int k = k0; k < k1; ++k) run += hist[k];
    int total;
    int off = block_exclusive_scan(run, &total);
    for (int k = k0; k < k1; ++k) {
        const int c = hist[k];
        cur[k] = off;
        starts[(long long)t * a.k2 + k] = b + off;
        off += c;
    }
    if (t == 0 && threadIdx.x == 0) starts[(long long)a.nTiles * a.k2] = tileStart[a.nTiles];
    __syncthreads();
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) {
        const int4 ent = tmp[i];
        entries[b + atomicAdd(&cur[key_of(ent.y)], 1)] = ent;
    }
}

// Work items: each (phase, tile) range of entries is cut into chunks of at most
// `chunk` entries.  Items are ordered phase-major so concurrently running
// workgroups read neighbouring C slices.  Item = {tile, first, count, exclusive}.
// The phases split the keys the entries use (from the cOffset range the
// binning noted in meta[4..5]) into equal runs, so a job that uses a part of
// its table (w-planes of a narrower range than the table spans) gets phases of
// that part's size.  Keys are floor(cOffset * k2 / cSize) by an integer
// division (one-level binning) or an FP64 product (two-level): both are
// monotone in cOffset, and the range covers what either gives at its ends.
__global__ __launch_bounds__(256) void k_plan_items(int pass, PlanDevArgs a,
                                                    const int *__restrict__ starts,
                                                    int *__restrict__ itemCount,
                                                    const int *__restrict__ itemStart,
                                                    int4 *__restrict__ items,
                                                    const int *__restrict__ meta)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= a.phases * a.nTiles) return;
    const int p = gid / a.nTiles, t = gid - p * a.nTiles;
    // XCD-major slot order: sub-range p goes to XCD p % 8, so each XCD's items
    // are contiguous and its L2 only sees its share of a phase's C slices
    const int id = a.xcd ? ((p & 7) * (a.phases >> 3) + (p >> 3)) * a.nTiles + t : gid;
    const int cLo = (int)(0x7fffffffu - (unsigned)meta[4]), cHi = meta[5] - 1;
    int kLo = 0, kHi = a.k2 - 1;
    if (cLo <= cHi) {
        const double scale = (double)a.k2 / (double)a.cSize;
        auto keyDiv = [&](int c) { return (int)(((long long)c * a.k2) / a.cSize); };
        auto keyMul = [&](int c) { return min((int)((double)c * scale), a.k2 - 1); };
        kLo = max(0, min(keyDiv(cLo), keyMul(cLo)));
        kHi = min(a.k2 - 1, max(keyDiv(cHi), keyMul(cHi)));
    }
    const long long span = kHi + 1 - kLo;
    const int k0 = kLo + (int)((long long)p * span / a.phases);
    const int k1 = kLo + (int)((long long)(p + 1) * span / a.phases);
    const long long b0 = (long long)t * a.k2;
    const int s = starts[b0 + k0], e = starts[b0 + k1];
    const int n = e - s;
    const int nc = (n + a.chunk - 1) / a.chunk;
    if (pass == 0) {
        itemCount[id] = nc;
    } else {
        const int excl = (a.phases == 1 && nc == 1) ? 1 : 0;
        const int o = itemStart[id];
        for (int k = 0; k < nc; ++k) {
            const int f = s + k * a.chunk;
            items[o + k] = make_int4(t, f, min(a.chunk, e - f), excl);
        }
    }
}

// meta: [0] items [1] queue head [2] entries [3] out of range [4..5] cOffset range
//       [8..15] per-XCD queue heads, [16..24] per-XCD item ranges (XCD-major plans)
__global__ void k_plan_meta(const int *__restrict__ itemStart, int nItemSlots,
                            const int *__restrict__ starts, long long nKeys, int xcdSlots,
                            int *__restrict__ meta)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        meta[0] = itemStart[nItemSlots];
        meta[1] = 0;
        meta[2] = starts[nKeys];
        for (int x = 0; x < 8; ++x) meta[8 + x] = 0;
        for (int x = 0; x <= 8; ++x)
            meta[16 + x] = xcdSlots ? itemStart[x * xcdSlots] : (x ? meta[0] : 0);
    }
}

__global__ void k_reset_queue(int *meta)
{
    if (threadIdx.x < 8) meta[8 + threadIdx.x] = 0;
    if (threadIdx.x == 0) meta[1] = 0;
}

// ---- tap storage ---------------------------------------------------------------
// FP32: complex64 taps (8 B).  FP16: (re, im) half pairs (4 B) scaled by the
// max-abs of their w-plane; accumulation stays FP32.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

template <bool HALF>
struct Taps;

template <>
struct Taps<false> {
    static constexpr unsigned kBytes = 8;
    __device__ static __forceinline__ float2 load(__amdgpu_buffer_rsrc_t r, unsigned voff)
    {
        return load_c(r, voff);
    }
};

template <>
struct Taps<true> {
    static constexpr unsigned kBytes = 4;
    __device__ static __forceinline__ float2 load(__amdgpu_buffer_rsrc_t r, unsigned voff)
    {
        const half2_t h = __builtin_bit_cast(half2_t, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
        return make_float2((float)h.x, (float)h.y);
    }
};

// Two consecutive taps (one 16-byte load for FP32, 8 bytes for FP16): the
// vector memory path moves data per group of four lanes, so wider loads carry
// twice the taps per unit of L1 throughput.
template <bool HALF>
__device__ __forceinline__ float4 load_pair(__amdgpu_buffer_rsrc_t r, unsigned voff)
{
    if constexpr (!HALF) {
        return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
    } else {
        const uint2 u = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
        const half2_t a = __builtin_bit_cast(half2_t, u.x), b = __builtin_bit_cast(half2_t, u.y);
        return make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
    }
}

// Four consecutive FP16 taps in one 16-byte load.
struct Quad {
    float2 t0, t1, t2, t3;
};

__device__ __forceinline__ float2 half_tap(unsigned w)
{
    const half2_t h = __builtin_bit_cast(half2_t, w);
    return make_float2((float)h.x, (float)h.y);
}

__device__ __forceinline__ Quad load_quad_half(__amdgpu_buffer_rsrc_t r, unsigned voff)
{
    const uint4 u = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
    return Quad{half_tap(u.x), half_tap(u.y), half_tap(u.z), half_tap(u.w)};
}

// a or b by a lane-dependent flag, as a bit blend: a plain select of two
// struct members can be rewritten by the compiler into an indexed load from
// the stack, which would put the quadruples in scratch memory
__device__ __forceinline__ float blend(bool pickB, float a, float b)
{
    const unsigned m = 0u - (unsigned)pickB;
    return __uint_as_float((__float_as_uint(a) & ~m) | (__float_as_uint(b) & m));
}

// tap i of a quadruple for a lane-dependent i in [0, 4)
__device__ __forceinline__ float2 pick4(const Quad &q, int i)
{
    const bool lo = i & 1, hi = i & 2;
    const float ax = blend(lo, q.t0.x, q.t1.x), ay = blend(lo, q.t0.y, q.t1.y);
    const float bx = blend(lo, q.t2.x, q.t3.x), by = blend(lo, q.t2.y, q.t3.y);
    return make_float2(blend(hi, ax, bx), blend(hi, ay, by));
}

__device__ __forceinline__ int entry_width(int w) { return w & 0xFFFF; }
__device__ __forceinline__ int entry_plane(const int4 &en, const JobDesc &d)
{
    return d.planeTaps > 0 ? en.y / d.planeTaps : (int)((unsigned)en.w >> 16);
}

template <bool HALF>
__device__ __forceinline__ float2 vis_value(const JobDesc &d, const int4 &en)
{
    if constexpr (!HALF) {
        if (!d.data) return make_float2(1.f, 0.f);   // PSF gridding: data == 1 + 0i
        return cload(static_cast<const float2 *>(d.data), en.z);
    } else {
        const half2_t v = cload(static_cast<const half2_t *>(d.data), en.z);
        const float sc = d.dataScale * cload(d.planeScale, entry_plane(en, d));
        return make_float2((float)v.x * sc, (float)v.y * sc);
    }
}

// Work-item source of the persistent kernels: one job (descriptor passed by
// value) or a batch (descriptors in device memory, located through the
// exclusive prefix sum of the jobs' item counts).
template <bool BATCH>
struct ItemSource {
    JobDesc one;
    const JobDesc *jobs;
    const int *prefix;
    int nJobs;
    int *queue;

    // Single job: items are XCD-major with per-XCD queues.  Batch: each XCD
    // walks its share of every job in job order (per-XCD prefix sums after the
    // job prefix, counters at queue[16 + 16 q]), then takes from the others.
    int xcd;
    int *bcast;   // one slot per workgroup (kernels that keep LDS for their tile)

    __device__ __forceinline__ int total() const
    {
        if constexpr (BATCH) return __builtin_amdgcn_readfirstlane(prefix[nJobs]);
        else return __builtin_amdgcn_readfirstlane(one.meta[0]);
    }
    // next work item (one lane calls this); >= total() when none is left
    __device__ __forceinline__ int acquire() const
    {
        if (!xcd) return atomicAdd(queue, 1);
        // own XCD first (hardware XCC_ID register), then steal from the others
        const int x0 = (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 7);
        if constexpr (BATCH) {
            const int n1 = nJobs + 1;
            for (int k = 0; k < 8; ++k) {
                const int q = (x0 + k) & 7;
                const int *px = prefix + n1 * (1 + q);
                const int tot = px[nJobs];
                if (tot == 0) continue;
                const int i = atomicAdd(queue + 16 + 16 * q, 1);
                if (i >= tot) continue;
                int lo = 0, hi = nJobs;   // px[lo] <= i < px[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (px[mid] <= i) lo = mid; else hi = mid;
                }
                return prefix[lo] + jobs[lo].meta[16 + q] + (i - px[lo]);
            }
            return 0x7fffffff;
        }
        for (int k = 0; k < 8; ++k) {
            const int q = (x0 + k) & 7;
            const int b = one.meta[16 + q], e = one.meta[17 + q];
            if (b >= e) continue;
            const int i = atomicAdd(one.meta + 8 + q, 1);
            if (b + i < e) return b + i;
        }
        return 0x7fffffff;
    }
    // item -> (job descriptor, local item index); wave-uniform
    __device__ __forceinline__ void locate(int item, JobDesc &d, int &local) const
    {
        if constexpr (BATCH) {
            int lo = 0, hi = nJobs;   // prefix[lo] <= item < prefix[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (__builtin_amdgcn_readfirstlane(prefix[mid]) <= item) lo = mid; else hi = mid;
            }
            d = jobs[lo];
            local = item - __builtin_amdgcn_readfirstlane(prefix[lo]);
        } else {
            d = one;
            local = item;
        }
    }
};

// Next work item for the whole workgroup without spending LDS: thread 0
// publishes the item it acquired during the previous one (`next`, taken from
// the queue one item ahead so that the queue's latency overlaps the work) in
// the workgroup's slot and acquires the following one.  Waiting for its memory
// counters completes the store in the CU's L2 before the barrier (a
// device-scope fence would also write back the L2), and the waves, all on the
// same CU, read the slot from that L2.  The caller reaches another barrier
// before thread 0 publishes again, so no wave can miss a value.  Every item
// acquired is processed: the loop ends on the first one past the end.
template <bool BATCH>
__device__ __forceinline__ int acquire_for_block(const ItemSource<BATCH> &src, int &next, int nItems)
{
    if (threadIdx.x == 0) {
        __hip_atomic_store(src.bcast + blockIdx.x, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);   // every counter of this wave drained
    }
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(src.bcast + blockIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (threadIdx.x == 0 && item < nItems) next = src.acquire();
    return item;
}

// The first item of a workgroup (thread 0 only; see acquire_for_block).
template <bool BATCH>
__device__ __forceinline__ int first_item(const ItemSource<BATCH> &src)
{
    return threadIdx.x == 0 ? src.acquire() : 0;
}

// The TW x TH grid tile at (Bx, By) into LDS (a workgroup of 256 threads),
// cells outside the grid as zero.  Every thread issues all its loads, two
// cells (16 bytes) each, before its first LDS write, so that their latencies
// overlap instead of adding up.
template <int TW, int TH, int NT = 256>
__device__ __forceinline__ void load_tile(float2 *tile, const float2 *__restrict__ grid, int gSize,
                                          int Bx, int By)
{
    constexpr int kPairs = TH * TW / 2;
    constexpr int kPer = (kPairs + NT - 1) / NT;
    constexpr int kBlk = kPer < 8 ? kPer : 8;   // loads in flight per thread (register bound)
#pragma unroll
    for (int i0 = 0; i0 < kPer; i0 += kBlk) {
        float4 v[kBlk];
#pragma unroll
        for (int b = 0; b < kBlk; ++b) {
            const int q = 2 * ((int)threadIdx.x + NT * (i0 + b));   // first cell of the pair
            const int row = By + q / TW, col = Bx + q % TW;
            v[b] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i0 + b < kPer && q < TH * TW && row < gSize) {
                const float2 *g = grid + (size_t)row * gSize + col;
                if (col + 1 < gSize) {
                    __builtin_memcpy(&v[b], g, 16);   // 8-byte aligned when gSize is odd
                } else if (col < gSize) {
                    const float2 a = *g;
                    v[b].x = a.x;
                    v[b].y = a.y;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < kBlk; ++b) {
            const int q = 2 * ((int)threadIdx.x + NT * (i0 + b));
            if (i0 + b < kPer && q < TH * TW) *reinterpret_cast<float4 *>(tile + q) = v[b];
        }
        // keep the next block's loads after these stores
        if constexpr (kPer > kBlk) __builtin_amdgcn_sched_barrier(0);
    }
}

// The same tile stored as two planes with a zero column on each side: tile
// column c is stored column j = c + 1 (j = 0 and j = TW + 1 hold zeros, set by
// zero_tile_pads), and stored column j sits in plane j & 1 at index
// row * (TW / 2 + 1) + j / 2.  A pair of adjacent cells is then two 8-byte
// reads, at one index of both planes or at consecutive ones, each contiguous
// over the lanes (no bank conflicts, where 16-byte pairs at 8-byte alignment
// read in two conflicting halves).
template <int TW, int TH, int NT = 256>
__device__ __forceinline__ void load_tile_planes(float2 *tile, const float2 *__restrict__ grid, int gSize,
                                                 int Bx, int By)
{
    constexpr int kPairs = TH * TW / 2;
    constexpr int RS = TW / 2 + 1;        // plane row stride (cells)
    constexpr int PC = TH * RS;           // cells a plane
    constexpr int kPer = (kPairs + NT - 1) / NT;
    constexpr int kBlk = kPer < 8 ? kPer : 8;
#pragma unroll
    for (int i0 = 0; i0 < kPer; i0 += kBlk) {
        float4 v[kBlk];
#pragma unroll
        for (int b = 0; b < kBlk; ++b) {
            const int q = 2 * ((int)threadIdx.x + NT * (i0 + b));
            const int row = By + q / TW, col = Bx + q % TW;
            v[b] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i0 + b < kPer && q < TH * TW && row < gSize) {
                const float2 *g = grid + (size_t)row * gSize + col;
                if (col + 1 < gSize) {
                    __builtin_memcpy(&v[b], g, 16);
                } else if (col < gSize) {
                    const float2 a = *g;
                    v[b].x = a.x;
                    v[b].y = a.y;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < kBlk; ++b) {
            const int q = 2 * ((int)threadIdx.x + NT * (i0 + b));
            if (i0 + b < kPer && q < TH * TW) {
                // cells (r, c) and (r, c + 1), c even: stored columns c + 1 (odd
                // plane) and c + 2 (even plane)
                const int i = (q / TW) * RS + (q % TW) / 2;
                tile[PC + i] = make_float2(v[b].x, v[b].y);
                tile[i + 1] = make_float2(v[b].z, v[b].w);
            }
        }
        if constexpr (kPer > kBlk) __builtin_amdgcn_sched_barrier(0);
    }
}

// The zero columns of load_tile_planes' layout (stored columns 0 and TW + 1);
// tile loads never write them.
template <int TW, int TH>
__device__ __forceinline__ void zero_tile_pads(float2 *tile)
{
    constexpr int RS = TW / 2 + 1, PC = TH * RS;
    for (int r = threadIdx.x; r < TH; r += blockDim.x) {
        tile[r * RS] = make_float2(0.f, 0.f);
        tile[PC + r * RS + TW / 2] = make_float2(0.f, 0.f);
    }
}

// ---- tiled gridder ------------------------------------------------------------
// attic/tConvolveMPI/Benchmark.cc:131-154 reorganised as owner-computes: a
// workgroup owns a TW x TH grid tile held in VGPRs.  Lane l covers column
// l % TW and, when TW < 64, row l / TW of each 64/TW-row group; wave w owns R
// consecutive row groups.  For every binned visibility the workgroup streams
// the part of its C slice that falls on the tile, accumulates, and writes the
// tile back once.  Persistent: workgroups pull items from a device queue.
template <int TW, int R, int NW, int RB, bool HALF, bool BATCH>
__global__ __launch_bounds__(NW * 64) void k_grid_tiled(ItemSource<BATCH> src)
{
    constexpr int RPI = 64 / TW;
    constexpr int TH = NW * R * RPI;
    constexpr unsigned TB = Taps<HALF>::kBytes;
    static_assert(R % RB == 0, "row block must divide the register rows");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lx = lane % TW, ly = lane / TW;
    __shared__ int s_item;
    const int nItems = src.total();
    const int q0 = wave * R;   // first row group of this wave

    int next = first_item(src);
    for (;;) {
        // `next` was acquired during the previous item: the queue's latency
        // overlaps the work
        if (threadIdx.x == 0) s_item = next;
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= nItems) break;
        if (threadIdx.x == 0) next = src.acquire();
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int4 *__restrict__ entries = jd.entries;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;
        const int e0 = it.y, e1 = it.y + it.z;

        float ar[R], ai[R];
#pragma unroll
        for (int j = 0; j < R; ++j) { ar[j] = 0.f; ai[j] = 0.f; }

        int4 enA = cload(entries, e0);
        int4 enB = cload(entries, min(e0 + 1, e1 - 1));
        float2 dA = vis_value<HALF>(jd, enA);
        for (int e = e0; e < e1; ++e) {
            const float2 dB = vis_value<HALF>(jd, enB);
            const int4 enC = cload(entries, min(e + 2, e1 - 1));

            const int u0 = enA.x & 0xFFFF, v0 = (int)((unsigned)enA.x >> 16);
            const int c0 = enA.y, s = entry_width(enA.w);
            const int cx = Bx + lx - u0;
            const int dy = By - v0;   // C row of tile row r is r + dy
            const unsigned lanev = ((unsigned)cx < (unsigned)s)
                                       ? (unsigned)(c0 + cx) * TB : kOOB;
#pragma unroll
            for (int jb = 0; jb < R; jb += RB) {
                const int r0 = (q0 + jb) * RPI;
                if (r0 + RB * RPI + dy > 0 && r0 + dy < s) {
                    float2 cv[RB];
#pragma unroll
                    for (int k = 0; k < RB; ++k) {
                        const int cy = r0 + k * RPI + ly + dy;
                        const unsigned vo = ((unsigned)cy < (unsigned)s)
                                                ? lanev + (unsigned)(cy * s) * TB : kOOB;
                        cv[k] = Taps<HALF>::load(rc, vo);
                    }
#pragma unroll
                    for (int k = 0; k < RB; ++k) {
                        ar[jb + k] = fmaf(dA.x, cv[k].x, fmaf(-dA.y, cv[k].y, ar[jb + k]));
                        ai[jb + k] = fmaf(dA.x, cv[k].y, fmaf(dA.y, cv[k].x, ai[jb + k]));
                    }
                }
            }
            enA = enB;
            dA = dB;
            enB = enC;
        }

        const int col = Bx + lx;
        const bool colOk = col < gSize;
        float2 *__restrict__ g0 = jd.grid + (size_t)(By + q0 * RPI + ly) * gSize + col;
        const int rowsLeft = gSize - (By + q0 * RPI + ly);   // rows j*RPI < rowsLeft exist
        if (it.w && !jd.sharedGrid) {
            // the tile is this item's alone: plain read-modify-write, every read
            // issued before the first write so that their latencies overlap
            float2 o[R];
#pragma unroll
            for (int j = 0; j < R; ++j)
                o[j] = (colOk && j * RPI < rowsLeft) ? g0[(size_t)j * RPI * gSize] : make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                if (colOk && j * RPI < rowsLeft && (ar[j] != 0.f || ai[j] != 0.f))
                    g0[(size_t)j * RPI * gSize] = make_float2(o[j].x + ar[j], o[j].y + ai[j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                if (colOk && j * RPI < rowsLeft && (ar[j] != 0.f || ai[j] != 0.f)) {
                    float *gp = reinterpret_cast<float *>(g0 + (size_t)j * RPI * gSize);
                    unsafeAtomicAdd(gp, ar[j]);
                    unsafeAtomicAdd(gp + 1, ai[j]);
                }
            }
        }
    }
}

// ---- tiled degridder ----------------------------------------------------------
// attic/tConvolveMPI/Benchmark.cc:165-193 reorganised: every wave holds the
// whole TW x TH grid tile in VGPRs (read once per work item); the waves of a
// workgroup split the tile's visibilities.  Each visibility's partial sum over
// the tile is reduced across the wave with DPP and added to the output with
// two FP32 atomics (the output is zeroed first: degridding overwrites).
template <int TW, int RD, int NW, int RB, bool HALF, bool BATCH>
__global__ __launch_bounds__(NW * 64) void k_degrid_tiled(ItemSource<BATCH> src)
{
    constexpr int RPI = 64 / TW;
    constexpr int TH = RD * RPI;
    constexpr unsigned TB = Taps<HALF>::kBytes;
    static_assert(RD % RB == 0, "row block must divide the register rows");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lx = lane % TW, ly = lane / TW;
    __shared__ int s_item;
    const int nItems = src.total();

    int next = first_item(src);
    for (;;) {
        // `next` was acquired during the previous item: the queue's latency
        // overlaps the work
        if (threadIdx.x == 0) s_item = next;
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= nItems) break;
        if (threadIdx.x == 0) next = src.acquire();
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int4 *__restrict__ entries = jd.entries;
        const float2 *__restrict__ grid = jd.grid;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;
        const int e0 = it.y, e1 = it.y + it.z;
        const int col = Bx + lx;

        float2 gv[RD];
#pragma unroll
        for (int j = 0; j < RD; ++j) {
            const int row = By + j * RPI + ly;
            gv[j] = (row < gSize && col < gSize) ? grid[(size_t)row * gSize + col]
                                                  : make_float2(0.f, 0.f);
        }

        const int first = e0 + wave;
        if (first >= e1) continue;
        int4 enA = cload(entries, first);
        for (int e = first; e < e1; e += NW) {
            const int4 enB = cload(entries, min(e + NW, e1 - 1));
            const int u0 = enA.x & 0xFFFF, v0 = (int)((unsigned)enA.x >> 16);
            const int c0 = enA.y, s = entry_width(enA.w);
            const int cx = Bx + lx - u0;
            const int dy = By - v0;
            const unsigned lanev = ((unsigned)cx < (unsigned)s)
                                       ? (unsigned)(c0 + cx) * TB : kOOB;
            float sr = 0.f, si = 0.f;
#pragma unroll
            for (int jb = 0; jb < RD; jb += RB) {
                const int r0 = jb * RPI;
                if (r0 + RB * RPI + dy > 0 && r0 + dy < s) {
                    float2 cv[RB];
#pragma unroll
                    for (int k = 0; k < RB; ++k) {
                        const int cy = r0 + k * RPI + ly + dy;
                        const unsigned vo = ((unsigned)cy < (unsigned)s)
                                                ? lanev + (unsigned)(cy * s) * TB : kOOB;
                        cv[k] = Taps<HALF>::load(rc, vo);
                    }
#pragma unroll
                    for (int k = 0; k < RB; ++k) {
                        sr = fmaf(gv[jb + k].x, cv[k].x, fmaf(-gv[jb + k].y, cv[k].y, sr));
                        si = fmaf(gv[jb + k].x, cv[k].y, fmaf(gv[jb + k].y, cv[k].x, si));
                    }
                }
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            if constexpr (HALF) {
                const float ps = cload(jd.planeScale, entry_plane(enA, jd));
                sr *= ps;
                si *= ps;
            }
            if (lane == 0) {
                float *op = reinterpret_cast<float *>(jd.out + enA.z);
                unsafeAtomicAdd(op, sr);
                unsafeAtomicAdd(op + 1, si);
            }
            enA = enB;
        }
    }
}

// ---- tiled gridder, two columns per lane ------------------------------------
// The register-tile gridder with each lane owning two adjacent tile columns,
// so every C row segment arrives in 16-byte loads (two taps per lane).  The
// tile's TW columns are cut into column groups of CGW; one load instruction
// covers a group's CGW columns over 128 / CGW tile rows (lane l: columns
// 2*(l % (CGW/2)) and +1 of the group, row l / (CGW/2) of the row group).  A
// load instruction costs the vector memory path the same whatever the number
// of its lanes inside the footprint (only one whose lanes are all out of range
// is cheap), so on footprint pieces narrower than the tile, narrow groups over
// several rows waste fewer lanes; groups a footprint does not reach are
// skipped.  A lane whose pair straddles the footprint's left or right edge
// loads the pair that starts (or ends) on the edge tap and uses one half, so
// every load stays inside the visibility's C slice.  Same per-cell arithmetic
// and entry order as the kernels above.
template <int TW, int CGW, int R, int NW, int RB, bool HALF, bool BATCH>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(4))) void k_grid_tiled2(ItemSource<BATCH> src)
{
    constexpr int LPR = CGW / 2;     // lanes per tile row of a column group
    constexpr int RPI = 64 / LPR;    // tile rows per instruction
    constexpr int NCG = TW / CGW;    // column groups
    constexpr int TH = NW * R * RPI;
    constexpr unsigned TB = Taps<HALF>::kBytes;
    static_assert(R % RB == 0, "row block must divide the register rows");
    static_assert(NCG * CGW == TW && LPR * RPI == 64, "column groups must tile a tile row");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lx = 2 * (lane % LPR), ly = lane / LPR;
    __shared__ int s_item;
    __shared__ float s_stage[NW][RPI * 2 * TW];   // one staged row group per wave
    const int nItems = src.total();
    const int q0 = wave * R;

    int next = first_item(src);
    for (;;) {
        // `next` was acquired during the previous item: the queue's latency
        // overlaps the work
        if (threadIdx.x == 0) s_item = next;
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= nItems) break;
        if (threadIdx.x == 0) next = src.acquire();
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int4 *__restrict__ entries = jd.entries;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;
        const int e0 = it.y, e1 = it.y + it.z;
        // Items whose entries may use the table's first or last slice (entries
        // are ordered by C-slice key, so only the first and last entries need
        // looking at) load every pair inside its slice (EDGE); the others read
        // straddling pairs whole.  Deciding per item rather than per entry keeps
        // the entry loop free of a branch whose two paths would make the
        // compiler copy the accumulators on every entry.
        const bool edgeItem = cload(entries, e0).y < jd.cEdgeLo || cload(entries, e1 - 1).y > jd.cEdgeHi;
        auto run = [&](auto edgeTag) {
        constexpr bool EDGE = decltype(edgeTag)::value;
        float ar0[NCG][R], ai0[NCG][R], ar1[NCG][R], ai1[NCG][R];
#pragma unroll
        for (int g = 0; g < NCG; ++g)
#pragma unroll
            for (int j = 0; j < R; ++j) { ar0[g][j] = ai0[g][j] = ar1[g][j] = ai1[g][j] = 0.f; }

        int4 enA = cload(entries, e0);
        int4 enB = cload(entries, min(e0 + 1, e1 - 1));
        float2 dA = vis_value<HALF>(jd, enA);
        for (int e = e0; e < e1; ++e) {
            const float2 dB = vis_value<HALF>(jd, enB);
            const int4 enC = cload(entries, min(e + 2, e1 - 1));
            const int u0 = enA.x & 0xFFFF, v0 = (int)((unsigned)enA.x >> 16);
            const int c0 = enA.y, s = entry_width(enA.w);
            const int dy = By - v0;
            const int cyl = ly + dy;       // C row of the lane's row in group 0, less r0
            const int sTB = s * (int)TB;
#pragma unroll
            for (int g = 0; g < NCG; ++g) {
                const int gx = Bx + g * CGW;   // the group's first grid column
                // a group the footprint's columns [u0, u0 + s) miss would only
                // load out of range (wave-uniform test)
                if (NCG > 1 && (gx + CGW <= u0 || gx >= u0 + s)) continue;
                const int cx = gx + lx - u0;   // C column of the lane's first cell
                const bool in0 = (unsigned)cx < (unsigned)s, in1 = (unsigned)(cx + 1) < (unsigned)s;
                if constexpr (!EDGE) {
                    // the pair (cx, cx + 1) even where it straddles the footprint's edge:
                    // the outside tap is the neighbouring one in the table, and its cell
                    // takes it with a zero visibility (an exact zero, tables are finite)
                    const float2 d0 = in0 ? dA : make_float2(0.f, 0.f);
                    const float2 d1 = in1 ? dA : make_float2(0.f, 0.f);
                    const unsigned lanev = (in0 || in1) ? (unsigned)(c0 + cx) * TB : kOOB;
#pragma unroll
                    for (int jb = 0; jb < R; jb += RB) {
                        const int r0 = (q0 + jb) * RPI;
                        if (r0 + RB * RPI + dy > 0 && r0 + dy < s) {
                            float4 cv[RB];
#pragma unroll
                            for (int k = 0; k < RB; ++k) {
                                const int cy = cyl + r0 + k * RPI;
                                cv[k] = load_pair<HALF>(
                                    rc, (unsigned)cy < (unsigned)s ? umad24(cy, sTB, lanev) : kOOB);
                            }
#pragma unroll
                            for (int k = 0; k < RB; ++k) {
                                float &r0v = ar0[g][jb + k], &i0v = ai0[g][jb + k];
                                float &r1v = ar1[g][jb + k], &i1v = ai1[g][jb + k];
                                r0v = fmaf(d0.x, cv[k].x, fmaf(-d0.y, cv[k].y, r0v));
                                i0v = fmaf(d0.x, cv[k].y, fmaf(d0.y, cv[k].x, i0v));
                                r1v = fmaf(d1.x, cv[k].z, fmaf(-d1.y, cv[k].w, r1v));
                                i1v = fmaf(d1.x, cv[k].w, fmaf(d1.y, cv[k].z, i1v));
                            }
                        }
                    }
                } else {
                    // slice at the table's start or end: load the pair that starts (or
                    // ends) on the edge tap, inside the slice, and pick the halves
                    const int start = min(max(cx, 0), s - 2);
                    const bool sh = cx > start;    // right edge: the wanted tap is the pair's second
                    const unsigned lanev = (in0 || in1) ? (unsigned)(c0 + start) * TB : kOOB;
#pragma unroll
                    for (int jb = 0; jb < R; jb += RB) {
                        const int r0 = (q0 + jb) * RPI;
                        if (r0 + RB * RPI + dy > 0 && r0 + dy < s) {
                            float4 cv[RB];
#pragma unroll
                            for (int k = 0; k < RB; ++k) {
                                const int cy = cyl + r0 + k * RPI;
                                cv[k] = load_pair<HALF>(
                                    rc, (unsigned)cy < (unsigned)s ? umad24(cy, sTB, lanev) : kOOB);
                            }
#pragma unroll
                            for (int k = 0; k < RB; ++k) {
                                // cell 0 takes tap 0 (tap 1 at the right edge); cell 1 takes
                                // tap 1 (tap 0 when the pair starts one column late)
                                float2 t0 = sh ? make_float2(cv[k].z, cv[k].w) : make_float2(cv[k].x, cv[k].y);
                                float2 t1 = in0 ? make_float2(cv[k].z, cv[k].w) : make_float2(cv[k].x, cv[k].y);
                                if (!in0) t0 = make_float2(0.f, 0.f);
                                if (!in1) t1 = make_float2(0.f, 0.f);
                                float &r0v = ar0[g][jb + k], &i0v = ai0[g][jb + k];
                                float &r1v = ar1[g][jb + k], &i1v = ai1[g][jb + k];
                                r0v = fmaf(dA.x, t0.x, fmaf(-dA.y, t0.y, r0v));
                                i0v = fmaf(dA.x, t0.y, fmaf(dA.y, t0.x, i0v));
                                r1v = fmaf(dA.x, t1.x, fmaf(-dA.y, t1.y, r1v));
                                i1v = fmaf(dA.x, t1.y, fmaf(dA.y, t1.x, i1v));
                            }
                        }
                    }
                }
            }
            enA = enB;
            dA = dB;
            enB = enC;
        }

        const int rowsLeft = gSize - (By + q0 * RPI + ly);
        if (it.w && !jd.sharedGrid) {
            // the tile is this item's alone: plain read-modify-write, every read
            // issued before the first write so that their latencies overlap
            // (RB rows at a time: the old values' registers)
#pragma unroll
            for (int g = 0; g < NCG; ++g) {
                const int col = Bx + g * CGW + lx;
                float2 *__restrict__ g0 = jd.grid + (size_t)(By + q0 * RPI + ly) * gSize + col;
                const bool c0ok = col < gSize, c1ok = col + 1 < gSize;
#pragma unroll
                for (int j0 = 0; j0 < R; j0 += RB) {
                    float2 o0[RB], o1[RB];
#pragma unroll
                    for (int b = 0; b < RB; ++b) {
                        const int j = j0 + b;
                        const float2 *gp = g0 + (size_t)j * RPI * gSize;
                        const bool rok = j * RPI < rowsLeft;
                        o0[b] = (rok && c0ok) ? gp[0] : make_float2(0.f, 0.f);
                        o1[b] = (rok && c1ok) ? gp[1] : make_float2(0.f, 0.f);
                    }
#pragma unroll
                    for (int b = 0; b < RB; ++b) {
                        const int j = j0 + b;
                        float2 *gp = g0 + (size_t)j * RPI * gSize;
                        const bool rok = j * RPI < rowsLeft;
                        if (rok && c0ok && (ar0[g][j] != 0.f || ai0[g][j] != 0.f))
                            gp[0] = make_float2(o0[b].x + ar0[g][j], o0[b].y + ai0[g][j]);
                        if (rok && c1ok && (ar1[g][j] != 0.f || ai1[g][j] != 0.f))
                            gp[1] = make_float2(o1[b].x + ar1[g][j], o1[b].y + ai1[g][j]);
                    }
                }
            }
        } else {
            // Atomic flush with every instruction on contiguous floats: each row
            // group goes through this wave's LDS staging rows so that lane l adds
            // float l (and l + 64, ...) of a tile row, instead of four floats at
            // a 16-byte stride per lane: four times fewer partly used cache lines
            // reach the atomic units.
            float *st = s_stage[wave];
            float *rowBase = reinterpret_cast<float *>(jd.grid + (size_t)(By + q0 * RPI) * gSize + Bx);
            const int rowsHere = gSize - (By + q0 * RPI);
            const int floatsHere = 2 * (gSize - Bx);
#pragma unroll
            for (int j = 0; j < R; ++j) {
#pragma unroll
                for (int g = 0; g < NCG; ++g)
                    *reinterpret_cast<float4 *>(st + ly * 2 * TW + 2 * (g * CGW + lx)) =
                        make_float4(ar0[g][j], ai0[g][j], ar1[g][j], ai1[g][j]);
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int k = 0; k < RPI * 2 * TW / 64; ++k) {
                    const int f = k * 64 + lane;              // float within the staged rows
                    const int q = f / (2 * TW), fc = f % (2 * TW);
                    const float v = st[f];
                    if (v != 0.f && j * RPI + q < rowsHere && fc < floatsHere)
                        unsafeAtomicAdd(rowBase + (size_t)(j * RPI + q) * gSize * 2 + fc, v);
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        };
        if (edgeItem) run(std::true_type{});
        else run(std::false_type{});
    }
}

// ---- FP16 storage: four taps per lane ------------------------------------
// With half (re, im) taps a 16-byte load carries four of them, so in FP16
// storage mode the register gridder gives each lane four adjacent tile columns
// and the LDS degridder walks footprints four columns at a time.  Edge lanes
// load the quadruple that starts (or ends) on the footprint's edge tap, so
// every load stays inside the visibility's C slice (kernel widths >= 4).
template <int TW, int R, int NW, int RB, bool BATCH>
__global__ __launch_bounds__(NW * 64) void k_grid_tiled4h(ItemSource<BATCH> src)
{
    constexpr int V = 4;
    constexpr int LPR = TW / V;      // lanes per tile row
    constexpr int RPI = 64 / LPR;    // tile rows per instruction
    constexpr int TH = NW * R * RPI;
    constexpr unsigned TB = 4;
    static_assert(R % RB == 0, "row block must divide the register rows");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lx = V * (lane % LPR), ly = lane / LPR;
    __shared__ int s_item;
    const int nItems = src.total();
    const int q0 = wave * R;

    int next = first_item(src);
    for (;;) {
        // `next` was acquired during the previous item: the queue's latency
        // overlaps the work
        if (threadIdx.x == 0) s_item = next;
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= nItems) break;
        if (threadIdx.x == 0) next = src.acquire();
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int4 *__restrict__ entries = jd.entries;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;
        const int e0 = it.y, e1 = it.y + it.z;

        float ar[R][V], ai[R][V];
#pragma unroll
        for (int j = 0; j < R; ++j)
#pragma unroll
            for (int c = 0; c < V; ++c) ar[j][c] = ai[j][c] = 0.f;

        int4 enA = cload(entries, e0);
        int4 enB = cload(entries, min(e0 + 1, e1 - 1));
        float2 dA = vis_value<true>(jd, enA);
        for (int e = e0; e < e1; ++e) {
            const float2 dB = vis_value<true>(jd, enB);
            const int4 enC = cload(entries, min(e + 2, e1 - 1));
            const int u0 = enA.x & 0xFFFF, v0 = (int)((unsigned)enA.x >> 16);
            const int c0 = enA.y, s = entry_width(enA.w);
            const int cx = Bx + lx - u0;               // C column of the lane's first cell
            const int start = min(max(cx, 0), s - V);  // the quadruple loaded
            const int off = cx - start;                // tap of cell c is off + c
            const unsigned lanev = (cx < s && cx + V > 0) ? (unsigned)(c0 + start) * TB : kOOB;
            bool ok[V];
#pragma unroll
            for (int c = 0; c < V; ++c) ok[c] = (unsigned)(cx + c) < (unsigned)s;
            const int dy = By - v0;
#pragma unroll
            for (int jb = 0; jb < R; jb += RB) {
                const int r0 = (q0 + jb) * RPI;
                if (r0 + RB * RPI + dy > 0 && r0 + dy < s) {
                    Quad t[RB];
#pragma unroll
                    for (int k = 0; k < RB; ++k) {
                        const int cy = r0 + k * RPI + ly + dy;
                        const unsigned vo = ((unsigned)cy < (unsigned)s)
                                                ? lanev + (unsigned)(cy * s) * TB : kOOB;
                        t[k] = load_quad_half(rc, vo);
                    }
#pragma unroll
                    for (int k = 0; k < RB; ++k)
#pragma unroll
                        for (int c = 0; c < V; ++c) {
                            float2 tc = pick4(t[k], off + c);
                            if (!ok[c]) tc = make_float2(0.f, 0.f);
                            ar[jb + k][c] = fmaf(dA.x, tc.x, fmaf(-dA.y, tc.y, ar[jb + k][c]));
                            ai[jb + k][c] = fmaf(dA.x, tc.y, fmaf(dA.y, tc.x, ai[jb + k][c]));
                        }
                }
            }
            enA = enB;
            dA = dB;
            enB = enC;
        }

        const int col = Bx + lx;
        const int rowsLeft = gSize - (By + q0 * RPI + ly);
        float2 *__restrict__ g0 = jd.grid + (size_t)(By + q0 * RPI + ly) * gSize + col;
        if (it.w && !jd.sharedGrid) {
#pragma unroll
            for (int j = 0; j < R; ++j)
#pragma unroll
                for (int c = 0; c < V; ++c) {
                    float2 *gp = g0 + (size_t)j * RPI * gSize + c;
                    if (j * RPI < rowsLeft && col + c < gSize && (ar[j][c] != 0.f || ai[j][c] != 0.f)) {
                        const float2 o = *gp;
                        *gp = make_float2(o.x + ar[j][c], o.y + ai[j][c]);
                    }
                }
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j)
#pragma unroll
                for (int c = 0; c < V; ++c) {
                    float *fp = reinterpret_cast<float *>(g0 + (size_t)j * RPI * gSize + c);
                    if (j * RPI < rowsLeft && col + c < gSize && (ar[j][c] != 0.f || ai[j][c] != 0.f)) {
                        unsafeAtomicAdd(fp, ar[j][c]);
                        unsafeAtomicAdd(fp + 1, ai[j][c]);
                    }
                }
        }
    }
}

// ---- tiled gridder / degridder, LDS tile with dense lanes --------------------
// The same work items as above, with the tile in LDS so that lanes map to the
// points of each visibility's footprint on the tile rather than to tile
// columns: an entry covering w x h points of the tile occupies ceil(w*h/64)
// lane groups instead of h full 64-lane rows (no lanes idle at the footprint's
// left and right edges).  Gridder: wave k owns the tile rows r with r % 4 == k
// (zero, accumulate and flush them), so no barrier or LDS atomic is needed and
// every cell accumulates the item's entries in entry order, rounded exactly as
// the register-tile gridder rounds them.
#ifndef TCV_LDS_U
#define TCV_LDS_U 8
#endif
constexpr int kLdsU = TCV_LDS_U;   // footprint points per lane whose taps load together

__device__ __forceinline__ void split_flat(int p, float rw, int w, int &row, int &col)
{
    // p < 2^12, w <= 64: (p + 0.5) / w is at least 1/(2w) away from an integer,
    // far beyond the rounding of the float reciprocal
    row = (int)(((float)p + 0.5f) * rw);
    col = p - umul24(row, w);
}

template <int TW, int TH, bool HALF, bool BATCH>
__global__ __launch_bounds__(256) void k_grid_lds(ItemSource<BATCH> src)
{
    constexpr unsigned TB = Taps<HALF>::kBytes;
    constexpr int OWN = (TH / 4) * TW;   // cells owned by one wave
    __shared__ float2 tile[TH * TW];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nItems = src.total();

    int next = first_item(src);
    for (;;) {
        const int item = acquire_for_block(src, next, nItems);
        __syncthreads();   // every wave has read the slot
        if (item >= nItems) break;
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int4 *__restrict__ entries = jd.entries;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;

        // own rows r = wave + 4 * (q / TW), column q % TW
        for (int q = lane; q < OWN; q += 64)
            tile[(wave + 4 * (q / TW)) * TW + (q % TW)] = make_float2(0.f, 0.f);

        for (int e = it.y; e < it.y + it.z; ++e) {
            const int4 en = cload(entries, e);
            const float2 d = vis_value<HALF>(jd, en);
            const int u0 = en.x & 0xFFFF, v0 = (int)((unsigned)en.x >> 16);
            const int c0 = en.y, s = entry_width(en.w);
            const int colLo = max(u0 - Bx, 0), colHi = min(u0 + s - Bx, TW);
            const int rowLo = max(v0 - By, 0), rowHi = min(v0 + s - By, TH);
            const int w = colHi - colLo;
            const int r0 = rowLo + ((wave - rowLo) & 3);   // first own row in range
            if (w <= 0 || r0 >= rowHi) continue;
            const int n = w * ((rowHi - r0 + 3) >> 2);
            const float rw = 1.0f / (float)w;
            const int cBase = c0 + (By + r0 - v0) * s + (Bx + colLo - u0);
            float2 *tBase = tile + r0 * TW + colLo;
            // kLdsU points per lane per pass: their taps are loaded together
            for (int p0 = lane; p0 < n; p0 += 64 * kLdsU) {
                float2 c[kLdsU];
                int to[kLdsU];
#pragma unroll
                for (int u = 0; u < kLdsU; ++u) {
                    const int p = p0 + 64 * u;
                    int rl, col;
                    split_flat(p, rw, w, rl, col);
                    to[u] = p < n ? 4 * rl * TW + col : -1;
                    c[u] = Taps<HALF>::load(rc, p < n ? (unsigned)(cBase + 4 * rl * s + col) * TB : kOOB);
                }
#pragma unroll
                for (int u = 0; u < kLdsU; ++u) {
                    if (to[u] >= 0) {
                        float2 *t = tBase + to[u];
                        float2 v = *t;
                        v.x = fmaf(d.x, c[u].x, fmaf(-d.y, c[u].y, v.x));
                        v.y = fmaf(d.x, c[u].y, fmaf(d.y, c[u].x, v.y));
                        *t = v;
                    }
                }
            }
        }

        for (int q = lane; q < OWN; q += 64) {
            const int r = wave + 4 * (q / TW), cc = q % TW;
            const int row = By + r, col = Bx + cc;
            const float2 v = tile[r * TW + cc];
            if (row < gSize && col < gSize && (v.x != 0.f || v.y != 0.f)) {
                float2 *gp = jd.grid + (size_t)row * gSize + col;
                if (it.w && !jd.sharedGrid) {
                    float2 o = *gp;
                    o.x += v.x;
                    o.y += v.y;
                    *gp = o;
                } else {
                    unsafeAtomicAdd(&gp->x, v.x);
                    unsafeAtomicAdd(&gp->y, v.y);
                }
            }
        }
    }
}

// Degridder with the tile in LDS: the workgroup loads the tile once per item,
// waves take whole entries, lanes walk the entry's footprint on the tile as a
// flat index, and the per-visibility sum is reduced with DPP as above.
template <int TW, int TH, bool HALF, bool BATCH>
__global__ __launch_bounds__(256) void k_degrid_lds(ItemSource<BATCH> src)
{
    constexpr unsigned TB = Taps<HALF>::kBytes;
    __shared__ float2 tile[TH * TW];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nItems = src.total();

    int next = first_item(src);
    for (;;) {
        // the barrier inside also means every wave is done with the previous
        // tile; the one after the tile load, that every wave has read the slot
        const int item = acquire_for_block(src, next, nItems);
        if (item >= nItems) break;
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int4 *__restrict__ entries = jd.entries;
        const float2 *__restrict__ grid = jd.grid;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;
        load_tile<TW, TH>(tile, grid, gSize, Bx, By);
        __syncthreads();

        for (int e = it.y + wave; e < it.y + it.z; e += 4) {
            const int4 en = cload(entries, e);
            const int u0 = en.x & 0xFFFF, v0 = (int)((unsigned)en.x >> 16);
            const int c0 = en.y, s = entry_width(en.w);
            const int colLo = max(u0 - Bx, 0), colHi = min(u0 + s - Bx, TW);
            const int rowLo = max(v0 - By, 0), rowHi = min(v0 + s - By, TH);
            const int w = colHi - colLo, h = rowHi - rowLo;
            if (w <= 0 || h <= 0) continue;
            const int n = w * h;
            const float rw = 1.0f / (float)w;
            const int cBase = c0 + (By + rowLo - v0) * s + (Bx + colLo - u0);
            const float2 *__restrict__ tBase = tile + rowLo * TW + colLo;
            float sr = 0.f, si = 0.f;
            for (int p0 = lane; p0 < n; p0 += 64 * kLdsU) {
                float2 c[kLdsU], g[kLdsU];
#pragma unroll
                for (int u = 0; u < kLdsU; ++u) {
                    const int p = p0 + 64 * u;
                    int rl, col;
                    split_flat(p, rw, w, rl, col);
                    c[u] = Taps<HALF>::load(rc, p < n ? (unsigned)(cBase + rl * s + col) * TB : kOOB);
                    g[u] = p < n ? tBase[rl * TW + col] : make_float2(0.f, 0.f);
                }
#pragma unroll
                for (int u = 0; u < kLdsU; ++u) {
                    sr = fmaf(g[u].x, c[u].x, fmaf(-g[u].y, c[u].y, sr));
                    si = fmaf(g[u].x, c[u].y, fmaf(g[u].y, c[u].x, si));
                }
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            if constexpr (HALF) {
                const float ps = cload(jd.planeScale, entry_plane(en, jd));
                sr *= ps;
                si *= ps;
            }
            if (lane == 0) {
                float *op = reinterpret_cast<float *>(jd.out + en.z);
                unsafeAtomicAdd(op, sr);
                unsafeAtomicAdd(op + 1, si);
            }
        }
    }
}

// Degridder with the LDS tile and tap pairs: lanes walk each entry's
// footprint on the tile densely, one pair of adjacent cells per lane and step
// (a 16-byte C load, two LDS cells).  The tile sits in LDS as two planes with
// a zero column on each side (load_tile_planes), so a pair is two
// conflict-free 8-byte reads; whether pairs start on an even or an odd stored
// column is chosen per entry (two compiled variants).  An odd width is walked
// as pairs over one more column whose cells count as zero: the tile's zero
// column where the footprint runs past the tile's edge (its tap is a tap of
// the footprint's own row), else a neighbouring column whose cells are
// masked (its taps are neighbours in the table: an exact zero contribution
// for a finite table).  Where such a tap could lie outside the table, the
// last column gets one pass with a lane per row instead.  Index arithmetic is
// 24-bit (full-rate), and only the last step of an entry checks for lanes
// past its end.
// NWD waves a workgroup: 8192-cell tiles (66.5 KB, two workgroups a CU) run eight when the
// job's C-slice phases stay near the L2's size (see degrid_tiled_t)
template <int TW, int TH, bool HALF, bool BATCH, int NWD = 4>
__global__ __launch_bounds__(NWD * 64) void k_degrid_lds2(ItemSource<BATCH> src)
{
    constexpr unsigned TB = Taps<HALF>::kBytes;
    constexpr int RS = TW / 2 + 1;                    // plane row stride (cells)
    constexpr unsigned PB = (unsigned)(TH * RS) * 8u; // bytes a plane
    __shared__ float2 tile[2 * TH * RS];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nItems = src.total();
    zero_tile_pads<TW, TH>(tile);   // ordered before any read by the barriers below

    int next = first_item(src);
    for (;;) {
        // the barrier inside also means every wave is done with the previous
        // tile; the one after the tile load, that every wave has read the slot
        const int item = acquire_for_block(src, next, nItems);
        if (item >= nItems) break;
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int nTaps = (int)(jd.cBytes / TB);
        const int4 *__restrict__ entries = jd.entries;
        const float2 *__restrict__ grid = jd.grid;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;
        load_tile_planes<TW, TH, NWD * 64>(tile, grid, gSize, Bx, By);
        __syncthreads();

        const int eEnd = it.y + it.z;
        for (int e = it.y + wave; e < eEnd; e += NWD) {
            const int4 en = cload(entries, e);
            const int u0 = en.x & 0xFFFF, v0 = (int)((unsigned)en.x >> 16);
            const int c0 = en.y, s = entry_width(en.w);
            const int colLo = max(u0 - Bx, 0), colHi = min(u0 + s - Bx, TW);
            const int rowLo = max(v0 - By, 0), rowHi = min(v0 + s - By, TH);
            const int w = colHi - colLo, h = rowHi - rowLo;
            if (w <= 0 || h <= 0) continue;
            // An odd width takes one more column, cells counted as zero:
            //  mode 3: the tile's zero column at the edge the footprint runs past
            //          (left: -1, right: TW); its taps lie in the footprint's rows
            //          unless the footprint ends exactly on that edge, when they
            //          are the neighbouring taps in the table
            //  mode 1: the next column (a real one, masked; its tap is the next
            //          one in the table)
            //  mode 2: the column before (masked; the tap before)
            //  mode 0: no added column: the last column gets a pass of its own
            //          (an added tap could lie outside the table)
            int mode = 0, colA = colLo, wA = w;
            if (w & 1) {
                const bool nextOk = c0 + s * s < nTaps, prevOk = c0 > 0;
                if (colLo == 0 && colHi < TW && (u0 < Bx || prevOk)) {
                    mode = 3;
                    colA = -1;
                    wA = w + 1;
                } else if (colHi == TW && colLo > 0 && (u0 + s > Bx + TW || nextOk)) {
                    mode = 3;
                    wA = w + 1;
                } else if (colHi < TW && nextOk) {
                    mode = 1;
                    wA = w + 1;
                } else if (colLo > 0 && prevOk) {
                    mode = 2;
                    colA = colLo - 1;
                    wA = w + 1;
                }
            }
            const int jA = colA + 1;                 // stored column of the walk's first cell
            const int wp = wA >> 1;                  // pairs per row
            const int kEdge = mode == 1 ? wp - 1 : 0;   // the pair with the masked cell
            const int n = wp * h;
            const float rw = 1.0f / (float)max(wp, 1);
            // byte offsets of cell (rowLo, colA): in C (its tap) and in the even
            // plane (index rowLo * RS + jA / 2; pair k of a row is k indices on)
            const unsigned cB = (unsigned)(c0 + (By + rowLo - v0) * s + Bx + colA - u0) * TB;
            const unsigned lB = (unsigned)(rowLo * RS + (jA >> 1)) * 8u;
            const int sB = s * (int)TB;
            float sr = 0.f, si = 0.f;
            // one step: pair k = p - rl * wp of row rl for p = p0 + lane + 64 u.
            // The offsets of the pair, base + rl * sB + 2 k TB (C) and
            // base + rl * RS * 8 + 8 k (plane), are written as
            // (base + 2 p TB) + rl * (sB - 2 wp TB): one 24-bit multiply-add each.
            // aL >= 0; aC < 0 when a kernel narrower than the tile gets its added column
            const int aC = sB - 2 * (int)TB * wp, aL = RS * 8 - 8 * wp;
            auto step = [&](int p0, auto tailTag, auto modeTag, auto oddTag, auto uTag) {
                constexpr bool TAIL = decltype(tailTag)::value;
                constexpr int MODE = decltype(modeTag)::value;
                constexpr bool ODD = decltype(oddTag)::value;   // pairs start on odd stored columns
                constexpr int U = decltype(uTag)::value;        // pairs a lane takes
                float4 c[U];
                float2 g0[U], g1[U];
                const int pl = p0 + lane;
                const float pf = (float)pl + 0.5f;
                const unsigned cP = cB + (unsigned)pl * (2u * TB), lP = lB + (unsigned)pl * 8u;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    // row = floor((p + 0.5) / wp): p + 0.5 is at least 1/(2 wp) from a
                    // multiple of wp, far beyond the rounding of the float arithmetic
                    const int rl = (int)fmaf(pf, rw, (float)(64 * u) * rw);
                    unsigned co = imad24(rl, aC, cP) + (unsigned)(64 * u) * (2u * TB);
                    unsigned lo = umad24(rl, aL, lP) + (unsigned)(64 * u) * 8u;
                    bool ok = true;
                    if constexpr (TAIL) {
                        ok = pl + 64 * u < n;
                        co = ok ? co : kOOB;
                        lo = ok ? lo : 0u;
                    }
                    c[u] = load_pair<HALF>(rc, co);
                    const char *t = reinterpret_cast<const char *>(tile) + lo;
                    if constexpr (ODD) {
                        g0[u] = *reinterpret_cast<const float2 *>(t + PB);
                        g1[u] = *reinterpret_cast<const float2 *>(t + 8);
                    } else {
                        g0[u] = *reinterpret_cast<const float2 *>(t);
                        g1[u] = *reinterpret_cast<const float2 *>(t + PB);
                    }
                    if constexpr (TAIL) {
                        if (!ok) g0[u] = g1[u] = make_float2(0.f, 0.f);
                    }
                    if constexpr (MODE == 1 || MODE == 2) {
                        // the pair on the added column: its outside cell takes zero
                        const bool edge = imad24(rl, -wp, pl) == kEdge - 64 * u;
                        if (MODE == 1 && edge) g1[u] = make_float2(0.f, 0.f);
                        if (MODE == 2 && edge) g0[u] = make_float2(0.f, 0.f);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    sr = fmaf(g0[u].x, c[u].x, fmaf(-g0[u].y, c[u].y, sr));
                    si = fmaf(g0[u].x, c[u].y, fmaf(g0[u].y, c[u].x, si));
                    sr = fmaf(g1[u].x, c[u].z, fmaf(-g1[u].y, c[u].w, sr));
                    si = fmaf(g1[u].x, c[u].w, fmaf(g1[u].y, c[u].z, si));
                }
            };
            auto walk = [&](auto modeTag, auto oddTag) {
                using Full = std::integral_constant<int, kLdsU>;
                int p0 = 0;
                for (; p0 + 64 * kLdsU <= n; p0 += 64 * kLdsU) step(p0, std::false_type{}, modeTag, oddTag, Full{});
                // the last pairs with only as many loads unrolled as they fill:
                // a narrow kernel's whole footprint is one such step, and the
                // idle lanes of unneeded loads cost their address arithmetic
                const int r = n - p0;
                if (r <= 0) return;
                if (r <= 64) step(p0, std::true_type{}, modeTag, oddTag, std::integral_constant<int, 1>{});
                else if (r <= 128 && kLdsU > 2)
                    step(p0, std::true_type{}, modeTag, oddTag, std::integral_constant<int, (kLdsU > 2 ? 2 : 1)>{});
                else if (r <= 256 && kLdsU > 4)
                    step(p0, std::true_type{}, modeTag, oddTag, std::integral_constant<int, (kLdsU > 4 ? 4 : 1)>{});
                else step(p0, std::true_type{}, modeTag, oddTag, Full{});
            };
            auto walkMode = [&](auto oddTag) {
                if (mode == 1) walk(std::integral_constant<int, 1>{}, oddTag);
                else if (mode == 2) walk(std::integral_constant<int, 2>{}, oddTag);
                else walk(std::integral_constant<int, 0>{}, oddTag);   // modes 0 and 3: nothing masked
            };
            if (jA & 1) walkMode(std::true_type{});
            else walkMode(std::false_type{});
            if ((w & 1) && mode == 0) {
                // the last column, a lane per row (64 rows a pass: tiles up to
                // 256 rows high)
                const unsigned col = (unsigned)(w - 1);
                const int j = colLo + w;   // its stored column
                const float2 *plane = tile + (j & 1) * (TH * RS) + (j >> 1);
                for (int r0 = 0; r0 < h; r0 += 64) {
                    const int r = r0 + lane;
                    const bool ok = r < h;
                    const float2 cv = Taps<HALF>::load(
                        rc, ok ? umad24(r, sB, cB + col * TB) : kOOB);
                    const float2 g = plane[min(rowLo + r, TH - 1) * RS];
                    const float2 gz = ok ? g : make_float2(0.f, 0.f);
                    sr = fmaf(gz.x, cv.x, fmaf(-gz.y, cv.y, sr));
                    si = fmaf(gz.x, cv.y, fmaf(gz.y, cv.x, si));
                }
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            if constexpr (HALF) {
                const float ps = cload(jd.planeScale, entry_plane(en, jd));
                sr *= ps;
                si *= ps;
            }
            if (lane == 0) {
                float *op = reinterpret_cast<float *>(jd.out + en.z);
                unsafeAtomicAdd(op, sr);
                unsafeAtomicAdd(op + 1, si);
            }
        }
    }
}

// FP16 degridder: LDS tile, footprints walked four columns at a time (one
// 16-byte load of four half taps per lane and group).
template <int TW, int TH, bool BATCH>
__global__ __launch_bounds__(256) void k_degrid_lds4h(ItemSource<BATCH> src)
{
    constexpr int V = 4;
    constexpr unsigned TB = 4;
    __shared__ float2 tile[TH * TW];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nItems = src.total();

    int next = first_item(src);
    for (;;) {
        // the barrier inside also means every wave is done with the previous
        // tile; the one after the tile load, that every wave has read the slot
        const int item = acquire_for_block(src, next, nItems);
        if (item >= nItems) break;
        JobDesc jd;
        int local;
        src.locate(item, jd, local);
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(jd.C, jd.cBytes);
        const int4 *__restrict__ entries = jd.entries;
        const float2 *__restrict__ grid = jd.grid;
        const int gSize = jd.gSize;
        const int4 it = cload(jd.items, local);
        const int tyi = it.x / jd.nTX, txi = it.x - tyi * jd.nTX;
        const int Bx = txi * TW, By = tyi * TH;
        load_tile<TW, TH>(tile, grid, gSize, Bx, By);
        __syncthreads();

        for (int e = it.y + wave; e < it.y + it.z; e += 4) {
            const int4 en = cload(entries, e);
            const int u0 = en.x & 0xFFFF, v0 = (int)((unsigned)en.x >> 16);
            const int c0 = en.y, s = entry_width(en.w);
            const int colLo = max(u0 - Bx, 0), colHi = min(u0 + s - Bx, TW);
            const int rowLo = max(v0 - By, 0), rowHi = min(v0 + s - By, TH);
            const int w = colHi - colLo, h = rowHi - rowLo;
            if (w <= 0 || h <= 0) continue;
            const int wq = (w + V - 1) / V;   // groups per row
            const int n = wq * h;
            const float rw = 1.0f / (float)wq;
            const int cxLo = Bx + colLo - u0;
            const int cRow = c0 + (By + rowLo - v0) * s;
            const float2 *tBase = tile + rowLo * TW + colLo;
            float sr = 0.f, si = 0.f;
            for (int p0 = lane; p0 < n; p0 += 64 * kLdsU) {
#pragma unroll
                for (int u = 0; u < kLdsU; ++u) {
                    const int p = p0 + 64 * u;
                    int rl, k;
                    split_flat(p, rw, wq, rl, k);
                    const int x = V * k;
                    const int cx = cxLo + x;
                    const int start = min(cx, s - V);   // s >= 4 for this kernel
                    const int off = cx - start;
                    const Quad t = load_quad_half(rc, p < n ? (unsigned)(cRow + rl * s + start) * TB : kOOB);
                    const float2 *tp = tBase + rl * TW + x;
#pragma unroll
                    for (int c = 0; c < V; ++c) {
                        const bool ok = p < n && x + c < w;
                        const float2 g = ok ? tp[c] : make_float2(0.f, 0.f);
                        const float2 tc = pick4(t, off + c);
                        sr = fmaf(g.x, tc.x, fmaf(-g.y, tc.y, sr));
                        si = fmaf(g.x, tc.y, fmaf(g.y, tc.x, si));
                    }
                }
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            const float ps = cload(jd.planeScale, entry_plane(en, jd));
            if (lane == 0) {
                float *op = reinterpret_cast<float *>(jd.out + en.z);
                unsafeAtomicAdd(op, sr * ps);
                unsafeAtomicAdd(op + 1, si * ps);
            }
        }
    }
}

// ---- FP16 storage conversion ----------------------------------------------------
// Per-plane max |component| of C, then taps / scale rounded to half.  Each thread
// walks a contiguous run of taps and flushes its running max with one atomicMax
// (positive floats order like their bit patterns) whenever the plane changes.
__device__ __forceinline__ int plane_of(long long t, const long long *planeStart, int nPlanes,
                                        long long planeTaps)
{
    if (planeTaps > 0) return (int)(t / planeTaps);
    int lo = 0, hi = nPlanes;   // planeStart[lo] <= t < planeStart[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (planeStart[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr int kRun = 1024;

__global__ __launch_bounds__(256) void k_plane_maxabs(const float2 *__restrict__ C, long long nTaps,
                                                      const long long *__restrict__ planeStart,
                                                      int nPlanes, long long planeTaps,
                                                      unsigned *__restrict__ maxbits)
{
    const long long t0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * kRun;
    if (t0 >= nTaps) return;
    const long long t1 = min(nTaps, t0 + kRun);
    int plane = plane_of(t0, planeStart, nPlanes, planeTaps);
    long long next = planeTaps > 0 ? (long long)(plane + 1) * planeTaps : planeStart[plane + 1];
    float m = 0.f;
    for (long long t = t0; t < t1; ++t) {
        if (t >= next) {
            atomicMax(&maxbits[plane], __float_as_uint(m));
            m = 0.f;
            plane = plane_of(t, planeStart, nPlanes, planeTaps);
            next = planeTaps > 0 ? (long long)(plane + 1) * planeTaps : planeStart[plane + 1];
        }
        const float2 c = C[t];
        m = fmaxf(m, fmaxf(fabsf(c.x), fabsf(c.y)));
    }
    atomicMax(&maxbits[plane], __float_as_uint(m));
}

__global__ __launch_bounds__(256) void k_taps_to_half(const float2 *__restrict__ C, long long nTaps,
                                                      const long long *__restrict__ planeStart,
                                                      int nPlanes, long long planeTaps,
                                                      unsigned *__restrict__ maxbits,
                                                      half2_t *__restrict__ out)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < nTaps; t += stride) {
        const int p = plane_of(t, planeStart, nPlanes, planeTaps);
        float sc = __uint_as_float(maxbits[p]);
        if (!(sc > 0.f)) sc = 1.f;
        const float2 c = C[t];
        half2_t hv;
        hv.x = (_Float16)(c.x / sc);
        hv.y = (_Float16)(c.y / sc);
        out[t] = hv;
    }
}

// max |component| over n complex values (one atomicMax per wave)
__global__ __launch_bounds__(256) void k_maxabs(const float2 *__restrict__ v, long long n,
                                                unsigned *__restrict__ maxbits)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    float m = 0.f;
    for (; i < n; i += stride) {
        const float2 c = v[i];
        m = fmaxf(m, fmaxf(fabsf(c.x), fabsf(c.y)));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, __float_as_uint(m));
}

__global__ __launch_bounds__(256) void k_vis_to_half(const float2 *__restrict__ v, long long n,
                                                     const unsigned *__restrict__ maxbits,
                                                     half2_t *__restrict__ out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    float sc = __uint_as_float(*maxbits);
    if (!(sc > 0.f)) sc = 1.f;
    for (; i < n; i += stride) {
        const float2 c = v[i];
        half2_t hv;
        hv.x = (_Float16)(c.x / sc);
        hv.y = (_Float16)(c.y / sc);
        out[i] = hv;
    }
}

int grid_blocks(size_t n, int bs)
{
    size_t b = (n + bs - 1) / bs;
    return (int)std::min<size_t>(b, 1u << 20);
}

}  // namespace

// ---- launchers ----------------------------------------------------------------

tcv_status launch_fill_u32(tcv_ctx *ctx, void *p, unsigned value, size_t n32)
{
    if (n32 == 0) return TCV_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_blocks(n32, 256)), dim3(256), 0, ctx->stream,
                       (unsigned *)p, value, n32);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_copy(tcv_ctx *ctx, void *dst, const void *src, size_t n16)
{
    if (n16 == 0) return TCV_OK;
    ProfScope ps(ctx, "copy");
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, (size_t)ctx->cus * 16);
    hipLaunchKernelGGL(k_copy_u4, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       (uint4 *)dst, (const uint4 *)src, n16);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_read_stream(tcv_ctx *ctx, const void *buf, size_t n16, int rounds, unsigned *sink)
{
    const unsigned blocks = (unsigned)std::max(1, ctx->cus) * 8u;
    const unsigned window = blocks * 256u;
    hipLaunchKernelGGL(k_read_stream, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4 *)buf,
                       (unsigned)(n16 - 1), window + 40961u, rounds, sink);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_offsets(tcv_ctx *ctx, const double *u, const double *v,
                          const double *w, int nSamples, const double *d_freq,
                          int nChan, double cellSize, double wCellSize, int wSize,
                          int gSize, int overSample, const int *d_sSize,
                          const int *d_cOffset0, int *iu, int *iv, int *wPlane,
                          int *cOffset, unsigned long long *d_numPerPlane,
                          unsigned long long *d_bad)
{
    const int total = nSamples * nChan;
    if (total == 0) return TCV_OK;
    const int useLds = (d_numPerPlane && wSize <= 8192) ? 1 : 0;
    const size_t lds = useLds ? (size_t)wSize * 4 : 0;
    ProfScope ps(ctx, "offsets");
    hipLaunchKernelGGL(k_offsets, dim3((total + 255) / 256), dim3(256), lds, ctx->stream,
                       u, v, w, nSamples, d_freq, nChan, cellSize, wCellSize, wSize,
                       gSize, overSample, d_sSize, d_cOffset0, iu, iv, wPlane, cOffset,
                       d_numPerPlane, d_bad, useLds);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_grid_plain(tcv_ctx *ctx, const GeomDev &g, int maxS,
                             const float2 *data, const float2 *C, float2 *grid)
{
    if (g.nVis == 0) return TCV_OK;
    ProfScope ps(ctx, "grid_plain");
    if (maxS * maxS <= 64)
        hipLaunchKernelGGL(k_grid_plain<64>, dim3(g.nVis), dim3(64), 0, ctx->stream, g, data,
                           C, grid, ctx->d_counters);
    else
        hipLaunchKernelGGL(k_grid_plain<256>, dim3(g.nVis), dim3(256), 0, ctx->stream, g, data,
                           C, grid, ctx->d_counters);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_degrid_plain(tcv_ctx *ctx, const GeomDev &g, const float2 *grid,
                               const float2 *C, float2 *out)
{
    if (g.nVis == 0) return TCV_OK;
    ProfScope ps(ctx, "degrid_plain");
    hipLaunchKernelGGL(k_degrid_plain, dim3((g.nVis + 3) / 4), dim3(256), 0, ctx->stream, g,
                       grid, C, out, ctx->d_counters);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_grid_plain_f64(tcv_ctx *ctx, const GeomDev &g, const float2 *data, const float2 *C,
                                 double2 *grid)
{
    if (g.nVis == 0) return TCV_OK;
    ProfScope ps(ctx, "grid_f64");
    hipLaunchKernelGGL(k_grid_plain_f64, dim3(g.nVis), dim3(256), 0, ctx->stream, g, data, C, grid,
                       ctx->d_counters);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_degrid_plain_f64(tcv_ctx *ctx, const GeomDev &g, const float2 *grid, const float2 *C,
                                   double2 *out)
{
    if (g.nVis == 0) return TCV_OK;
    ProfScope ps(ctx, "degrid_f64");
    hipLaunchKernelGGL(k_degrid_plain_f64, dim3((g.nVis + 3) / 4), dim3(256), 0, ctx->stream, g, grid,
                       C, out, ctx->d_counters);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

long long scan_tmp_len(long long n)
{
    // sums + scanned sums for every level
    long long len = 0;
    while (true) {
        const long long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
        if (nb <= 1) break;
        len += 2 * (nb + 1);
        n = nb;
    }
    return len + 2;
}

static tcv_status scan_exclusive(tcv_ctx *ctx, const int *in, int *out, long long n, int *tmp)
{
    const long long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_BS), 0, ctx->stream, in, n,
                           (const int *)nullptr, out);
        TCV_HIP(ctx, hipGetLastError());
        return TCV_OK;
    }
    int *sums = tmp;
    int *sumsScan = tmp + nb + 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_BS), 0, ctx->stream, in, n,
                       sums);
    TCV_HIP(ctx, hipGetLastError());
    tcv_status st = scan_exclusive(ctx, sums, sumsScan, nb, tmp + 2 * (nb + 1));
    if (st != TCV_OK) return st;
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_BS), 0, ctx->stream, in, n,
                       (const int *)sumsScan, out);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status plan_build_kernels(tcv_ctx *ctx, tcv_plan *p, const GeomDev &g)
{
    const PlanShape &sh = p->sh;
    PlanDevArgs a{sh.tw, sh.th, sh.nTX, sh.k2, sh.phases, sh.chunk, sh.nTiles, sh.xcd, g.cSize};
    ProfScope ps(ctx, "plan");
    tcv_status st;
    if ((st = launch_fill_u32(ctx, p->counts, 0, (size_t)sh.nKeys + 1)) != TCV_OK) return st;
    if ((st = launch_fill_u32(ctx, p->meta, 0, 32)) != TCV_OK) return st;
    const bool lds = sh.nKeys <= kLdsKeys;
    const int nb = lds ? (g.nVis + 256 * kBinVPT - 1) / (256 * kBinVPT) : (g.nVis + 255) / 256;
    auto bin = [&](int pass) -> tcv_status {
        if (g.nVis <= 0) return TCV_OK;
        if (lds)
            hipLaunchKernelGGL(k_plan_bin_lds, dim3(nb), dim3(256),
                               (size_t)(pass ? 2 : 1) * sh.nKeys * sizeof(int), ctx->stream, pass,
                               g, a, p->counts, p->cursor, p->entries, p->meta);
        else
            hipLaunchKernelGGL(k_plan_bin, dim3(nb), dim3(256), 0, ctx->stream, pass, g, a,
                               p->counts, p->cursor, p->entries, p->meta);
        TCV_HIP(ctx, hipGetLastError());
        return TCV_OK;
    };
    if (p->entriesTmp) {
        // Two levels: bin by tile alone (the tile histogram fits LDS, so each
        // block touches each global tile counter at most once), then order
        // every tile's entries by C-slice key in its own workgroup.  Global
        // atomics drop from one per entry to one per (block, tile).
        PlanDevArgs t1 = a;
        t1.k2 = 1;
        int *tileCount = p->counts, *tileStart = p->cursor, *tileCursor = p->cursor + sh.nTiles + 1;
        const int nb1 = (g.nVis + 256 * kBinVPT - 1) / (256 * kBinVPT);
        if (g.nVis > 0) {
            hipLaunchKernelGGL(k_plan_bin_lds, dim3(nb1), dim3(256), (size_t)sh.nTiles * sizeof(int),
                               ctx->stream, 0, g, t1, tileCount, tileCursor, p->entriesTmp, p->meta);
            TCV_HIP(ctx, hipGetLastError());
        }
        if ((st = scan_exclusive(ctx, tileCount, tileStart, sh.nTiles, p->scanTmp)) != TCV_OK) return st;
        TCV_HIP(ctx, hipMemcpyAsync(tileCursor, tileStart, sizeof(int) * sh.nTiles,
                                    hipMemcpyDeviceToDevice, ctx->stream));
        if (g.nVis > 0) {
            hipLaunchKernelGGL(k_plan_bin_lds, dim3(nb1), dim3(256), (size_t)2 * sh.nTiles * sizeof(int),
                               ctx->stream, 1, g, t1, tileCount, tileCursor, p->entriesTmp, p->meta);
            TCV_HIP(ctx, hipGetLastError());
        }
        hipLaunchKernelGGL(k_plan_tile_keys, dim3(sh.nTiles), dim3(256), (size_t)2 * sh.k2 * sizeof(int),
                           ctx->stream, a, (const int *)tileStart, (const int4 *)p->entriesTmp,
                           p->entries, p->starts);
        TCV_HIP(ctx, hipGetLastError());
    } else {
        if ((st = bin(0)) != TCV_OK) return st;
        if ((st = scan_exclusive(ctx, p->counts, p->starts, sh.nKeys, p->scanTmp)) != TCV_OK)
            return st;
        TCV_HIP(ctx, hipMemcpyAsync(p->cursor, p->starts, sizeof(int) * sh.nKeys,
                                    hipMemcpyDeviceToDevice, ctx->stream));
        if ((st = bin(1)) != TCV_OK) return st;
    }
    // work items: cuts of the sorted entries into (phase, tile) ranges
    auto items = [&](int phases, int xcd, int *itemCount, int *itemStart, int4 *itemsOut,
                     int *meta) -> tcv_status {
        PlanDevArgs b = a;
        b.phases = phases;
        b.xcd = xcd;
        const int slots = phases * sh.nTiles;
        hipLaunchKernelGGL(k_plan_items, dim3((slots + 255) / 256), dim3(256), 0, ctx->stream, 0,
                           b, p->starts, itemCount, (const int *)itemStart, itemsOut,
                           (const int *)p->meta);
        TCV_HIP(ctx, hipGetLastError());
        tcv_status s2 = scan_exclusive(ctx, itemCount, itemStart, slots, p->scanTmp);
        if (s2 != TCV_OK) return s2;
        hipLaunchKernelGGL(k_plan_items, dim3((slots + 255) / 256), dim3(256), 0, ctx->stream, 1,
                           b, p->starts, itemCount, (const int *)itemStart, itemsOut,
                           (const int *)p->meta);
        TCV_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_plan_meta, dim3(1), dim3(64), 0, ctx->stream, (const int *)itemStart,
                           slots, (const int *)p->starts, sh.nKeys,
                           xcd ? (phases / 8) * sh.nTiles : 0, meta);
        TCV_HIP(ctx, hipGetLastError());
        return TCV_OK;
    };
    if ((st = items(sh.phases, sh.xcd, p->itemCount, p->itemStart, p->items, p->meta)) != TCV_OK)
        return st;
    if (p->metaG != p->meta) {
        // the out-of-range count lives in meta[3]; the gridder's meta gets its own copy
        if ((st = launch_fill_u32(ctx, p->metaG, 0, 32)) != TCV_OK) return st;
        if ((st = items(sh.phasesG, sh.xcdG, p->itemCountG, p->itemStartG, p->itemsG, p->metaG)) != TCV_OK)
            return st;
    }
    return TCV_OK;
}

// ---- tiled launch dispatch ----------------------------------------------------

namespace {

constexpr int kNW = 4;   // waves per workgroup
#ifndef TCV_RB
#define TCV_RB 4
#endif
constexpr int kRB = TCV_RB;   // row groups per uniform range test (loads issued together)
// the pair gridder issues all its row loads of an entry together (measured 2 %
// faster than four at a time on the spectral workload)
#ifndef TCV_RB2
#define TCV_RB2 8
#endif
constexpr int kRB2 = TCV_RB2;
// tiles taller than 64 rows: row groups per range test of the pair gridder
// (a wave's groups there hold several tile rows each)
#ifndef TCV_RB2_TALL
#define TCV_RB2_TALL 8
#endif
constexpr int kRB2Tall = TCV_RB2_TALL;

template <int TW, int TH>
struct TileCfg {
    // tiles wider than a wave or taller than 64 rows have only the pair
    // kernels (kernel 3)
    static constexpr bool kPairOnly = TW > 64 || TH > 64;
    static constexpr int RB2 = TH > 64 ? kRB2Tall : kRB2;
    static constexpr int RPI = TW > 64 ? 1 : 64 / TW;
    static constexpr int RG = kPairOnly ? 0 : TH / (kNW * RPI);   // gridder rows groups per wave
    static constexpr int RD = kPairOnly ? 0 : TH / RPI;           // degridder row groups per wave
    static constexpr int RG2 = TH / (kNW * (128 / TW));   // two-column gridder (0: not built)
    static constexpr int RG4 = TH * TW / (kNW * 256);      // four-column FP16 gridder (0: not built)
};

int persistent_blocks(tcv_ctx *ctx, const void *kernel, int bs)
{
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, bs, 0) != hipSuccess || per < 1)
        per = 1;
    // every block has a hand-over slot in the context
    return std::min(std::max(1, ctx->cus) * per, kMaxPersistentBlocks);
}

// Item prefix over the jobs (thread 0) and, after it, one prefix per XCD over
// the jobs' per-XCD item ranges (threads 1..8); resets the queue counters.
__global__ void k_batch_prefix(const JobDesc *__restrict__ jobs, int nJobs, int *__restrict__ prefix,
                               int *__restrict__ queue)
{
    const int t = threadIdx.x;
    if (blockIdx.x != 0 || t > 8) return;
    int *out = prefix + (nJobs + 1) * t;
    int run = 0;
    for (int j = 0; j < nJobs; ++j) {
        out[j] = run;
        const int *m = jobs[j].meta;
        run += t == 0 ? m[0] : m[16 + t] - m[15 + t];
    }
    out[nJobs] = run;
    queue[16 * t] = 0;   // t = 0: the single queue; t = 1 + q: XCD q's counter
}

// the pair gridder over column groups of CGW (TH gives every wave whole row groups)
template <int TW, int TH, int CGW, bool HALF, bool BATCH>
auto grid2_kernel() -> void (*)(ItemSource<BATCH>)
{
    constexpr int R = TH / (kNW * (128 / CGW));
    constexpr int RB = TileCfg<TW, TH>::RB2;
    return k_grid_tiled2<TW, CGW, R, kNW, (R % RB == 0 ? RB : R), HALF, BATCH>;
}

template <int TW, int TH, bool HALF, bool BATCH>
tcv_status grid_tiled_t(tcv_ctx *ctx, const ItemSource<BATCH> &src, int kern_kind, int cols)
{
    using T = TileCfg<TW, TH>;
    void (*kern)(ItemSource<BATCH>) = nullptr;
    if constexpr (!T::kPairOnly)
        kern = kern_kind == 2 ? k_grid_lds<TW, TH, HALF, BATCH>
                              : k_grid_tiled<TW, T::RG, kNW, (T::RG % kRB == 0 ? kRB : T::RG), HALF, BATCH>;
    if constexpr (T::RG2 >= 1) {
        if (kern_kind >= 3 || T::kPairOnly) {
            kern = grid2_kernel<TW, TH, TW, HALF, BATCH>();
            if constexpr (TW >= 64 && TH >= kNW * 256 / TW)
                if (cols == TW / 2) kern = grid2_kernel<TW, TH, TW / 2, HALF, BATCH>();
            if constexpr (TW >= 64 && TH >= kNW * 512 / TW)
                if (cols == TW / 4) kern = grid2_kernel<TW, TH, TW / 4, HALF, BATCH>();
        }
    }
    if constexpr (HALF && T::RG4 >= 1 && !T::kPairOnly) {
        // FP16 storage: four taps per 16-byte load
        if (kern_kind == 4)
            kern = k_grid_tiled4h<TW, T::RG4, kNW, (T::RG4 % kRB == 0 ? kRB : T::RG4), BATCH>;
    }
    const int blocks = persistent_blocks(ctx, (const void *)kern, kNW * 64);
    ProfScope ps(ctx, BATCH ? "grid_batch" : (HALF ? "grid_tiled_half" : "grid_tiled"));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kNW * 64), 0, ctx->stream, src);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

// manyWaves: the pair degridder on 8192-cell tiles runs eight-wave workgroups.
// Measured on 128 x 64 tiles: 16 % faster with 3.8 MB phases (spectral
// workload) and on the uneven batch, 22 % slower with 70 MB phases (a 160 000
// visibility job cut into 4 phases), where more waves only add misses.
template <int TW, int TH, bool HALF, bool BATCH>
tcv_status degrid_tiled_t(tcv_ctx *ctx, const ItemSource<BATCH> &src, int kern_kind, bool manyWaves)
{
    using T = TileCfg<TW, TH>;
    void (*kern)(ItemSource<BATCH>) = k_degrid_lds2<TW, TH, HALF, BATCH>;
    int threads = kNW * 64;
    if constexpr (TW * TH >= 8192) {
        if (manyWaves) {
            kern = k_degrid_lds2<TW, TH, HALF, BATCH, 8>;
            threads = 8 * 64;
        }
    }
    if constexpr (!T::kPairOnly) {
        if (kern_kind == 2) kern = k_degrid_lds<TW, TH, HALF, BATCH>;
        else if (kern_kind < 3)
            kern = k_degrid_tiled<TW, T::RD, kNW, (T::RD % kRB == 0 ? kRB : T::RD), HALF, BATCH>;
        if (kern_kind < 3) threads = kNW * 64;
    }
    if constexpr (HALF && !T::kPairOnly) {
        if (kern_kind == 4) {
            kern = k_degrid_lds4h<TW, TH, BATCH>;   // four FP16 taps per load
            threads = kNW * 64;
        }
    }
    const int blocks = persistent_blocks(ctx, (const void *)kern, threads);
    ProfScope ps(ctx, BATCH ? "degrid_batch" : (HALF ? "degrid_tiled_half" : "degrid_tiled"));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), 0, ctx->stream, src);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

}  // namespace

#define TCV_TILE_CONFIGS(X) \
    X(128, 64) X(64, 128) X(32, 256) X(64, 64) X(64, 32) X(32, 64) X(32, 32) X(16, 64) X(16, 32) \
    X(8, 64) X(8, 32)

bool plan_two_level(const PlanShape &sh)
{
    return sh.k2 >= 2 && sh.k2 <= 8192 && sh.nTiles <= kLdsKeys;
}

bool grid2_supported(int tw, int th)
{
    return th >= kNW * (128 / tw);
}

bool grid_cols_supported(int tw, int th, int cols)
{
    if (cols == tw) return true;
    if (tw < 64 || cols < 16 || (cols != tw / 2 && cols != tw / 4)) return false;
    return th >= kNW * 128 / cols && tiled_config_supported(tw, th);
}

bool tiled_config_supported(int tw, int th)
{
#define X(a, b) if (tw == a && th == b) return true;
    TCV_TILE_CONFIGS(X)
#undef X
    return false;
}

JobDesc make_job_desc(const tcv_plan *p, bool forGrid, const void *data, const void *C, float2 *grid,
                      float2 *out, const HalfStorage *half)
{
    JobDesc d{};
    d.items = forGrid ? p->itemsG : p->items;
    d.entries = p->entries;
    d.meta = forGrid ? p->metaG : p->meta;
    d.data = data;
    d.C = C;
    d.grid = grid;
    d.out = out;
    d.planeScale = half ? half->planeScale : nullptr;
    d.cBytes = (unsigned)(p->cSize * (half ? 4 : 8));
    d.nTX = p->sh.nTX;
    d.gSize = p->gSize;
    d.planeTaps = half ? half->planeTaps : 0;
    d.dataScale = half ? half->dataScale : 1.f;
    d.nVis = p->nVis;
    // one C-slice key spans cSize / k2 taps; two more for the rounding of the key
    const long long keyTaps = (p->cSize + p->sh.k2 - 1) / p->sh.k2 + 2;
    d.cEdgeLo = (int)std::min<long long>(keyTaps, 0x7fffffffLL);
    d.cEdgeHi = (int)std::max<long long>(p->cSize - (long long)p->sh.maxS * p->sh.maxS - keyTaps, -1);
    return d;
}

tcv_status launch_grid_tiled(tcv_ctx *ctx, tcv_plan *p, const void *data, const void *C,
                             float2 *grid, const HalfStorage *half)
{
    ItemSource<false> src{make_job_desc(p, true, data, C, grid, nullptr, half), nullptr, nullptr, 0,
                          p->metaG + 1, p->sh.xcdG, ctx->d_bcast};
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, ctx->stream, p->metaG);
    TCV_HIP(ctx, hipGetLastError());
    // the FP16 four-tap gridder needs kernel widths of at least 4
    const int kern = (half && p->sh.kern == 4 && p->sh.minS < 4) ? 3 : p->sh.kern;
#define X(a, b)                                                                        \
    if (p->sh.tw == a && p->sh.th == b)                                                \
        return half ? grid_tiled_t<a, b, true, false>(ctx, src, kern, p->sh.gridCols)  \
                    : grid_tiled_t<a, b, false, false>(ctx, src, kern, p->sh.gridCols);
    TCV_TILE_CONFIGS(X)
#undef X
    return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tile shape not built");
}

// C-slice phase size up to which the 8192-cell-tile degridder runs eight waves
constexpr double kManyWavesPhaseBytes = 32.0 * (1 << 20);

tcv_status launch_degrid_tiled(tcv_ctx *ctx, tcv_plan *p, const float2 *grid, const void *C,
                               float2 *out, const HalfStorage *half)
{
    ItemSource<false> src{make_job_desc(p, false, nullptr, C, const_cast<float2 *>(grid), out, half),
                          nullptr, nullptr, 0, p->meta + 1, p->sh.xcd, ctx->d_bcast};
    tcv_status st = launch_fill_u32(ctx, out, 0, (size_t)p->nVis * 2);
    if (st != TCV_OK) return st;
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, ctx->stream, p->meta);
    TCV_HIP(ctx, hipGetLastError());
    // the FP16 four-tap degridder needs kernel widths of at least 4
    const int kern = (half && p->sh.dkern == 4 && p->sh.minS < 4) ? 3 : p->sh.dkern;
    const double phaseBytes = (double)p->cSize * (half ? 4 : 8) / std::max(1, p->sh.phases);
    const bool many = phaseBytes <= kManyWavesPhaseBytes;
#define X(a, b)                                                                        \
    if (p->sh.tw == a && p->sh.th == b)                                                \
        return half ? degrid_tiled_t<a, b, true, false>(ctx, src, kern, many)          \
                    : degrid_tiled_t<a, b, false, false>(ctx, src, kern, many);
    TCV_TILE_CONFIGS(X)
#undef X
    return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tile shape not built");
}

tcv_status launch_batch(tcv_ctx *ctx, tcv_batch *b, bool degrid)
{
    if (degrid)
        for (const JobDesc &d : b->hostD) {
            tcv_status st = launch_fill_u32(ctx, d.out, 0, (size_t)d.nVis * 2);
            if (st != TCV_OK) return st;
        }
    hipLaunchKernelGGL(k_batch_prefix, dim3(1), dim3(64), 0, ctx->stream,
                       (const JobDesc *)(degrid ? b->d_jobsD : b->d_jobs), b->nJobs, b->d_prefix,
                       b->d_queue);
    TCV_HIP(ctx, hipGetLastError());
    ItemSource<true> src{JobDesc{}, degrid ? b->d_jobsD : b->d_jobs, b->d_prefix, b->nJobs,
                         b->d_queue, 1, ctx->d_bcast};
    // batches run FP32 storage
#define X(a, c)                                                                        \
    if (b->tw == a && b->th == c)                                                      \
        return degrid ? degrid_tiled_t<a, c, false, true>(ctx, src, b->dkern, true)    \
                      : grid_tiled_t<a, c, false, true>(ctx, src, b->kern, b->gridCols);
    TCV_TILE_CONFIGS(X)
#undef X
    return tcv_fail(ctx, TCV_ERR_UNSUPPORTED, "tile shape not built");
}

tcv_status launch_ctable_to_half(tcv_ctx *ctx, const float2 *C, long long nTaps, int nPlanes,
                                 const long long *d_planeStart, long long planeTaps,
                                 unsigned *d_maxbits, void *out)
{
    if (nTaps <= 0) return TCV_OK;
    tcv_status st = launch_fill_u32(ctx, d_maxbits, 0, (size_t)nPlanes);
    if (st != TCV_OK) return st;
    const long long runs = (nTaps + kRun - 1) / kRun;
    hipLaunchKernelGGL(k_plane_maxabs, dim3((unsigned)((runs + 255) / 256)), dim3(256), 0,
                       ctx->stream, C, nTaps, d_planeStart, nPlanes, planeTaps, d_maxbits);
    TCV_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_taps_to_half, dim3(grid_blocks((size_t)nTaps, 256)), dim3(256), 0,
                       ctx->stream, C, nTaps, d_planeStart, nPlanes, planeTaps, d_maxbits,
                       (half2_t *)out);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

tcv_status launch_vis_to_half(tcv_ctx *ctx, const float2 *data, long long n, unsigned *d_maxbits,
                              void *out)
{
    tcv_status st = launch_fill_u32(ctx, d_maxbits, 0, 1);
    if (st != TCV_OK || n <= 0) return st;
    hipLaunchKernelGGL(k_maxabs, dim3(std::min(grid_blocks((size_t)n, 256), 1024)), dim3(256), 0,
                       ctx->stream, data, n, d_maxbits);
    TCV_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_vis_to_half, dim3(grid_blocks((size_t)n, 256)), dim3(256), 0, ctx->stream,
                       data, n, (const unsigned *)d_maxbits, (half2_t *)out);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}

// ---- attic C(w) taps on the device ----------------------------------------------
// attic/tConvolveMPI/Benchmark.cc:236-269 evaluated per tap in FP64 on the GPU.
// The normalisation divides by the FP64 sum of |C| (the reference accumulates
// that sum in float, which stops growing near 2^24 on large tables); use the
// host tcv_attic_initC for tables bit-identical to the reference.
namespace {

__global__ __launch_bounds__(256) void k_attic_taps(double freq0, double cellSize, int wSize, int s,
                                                    int os, double wCellSize, float2 *__restrict__ C,
                                                    double *__restrict__ sum)
{
    const long long n = (long long)s * s * os * os * wSize;
    const int cc = (s - 1) / 2;
    double acc = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long q = t;
        const int i = (int)(q % s); q /= s;
        const int j = (int)(q % s); q /= s;
        const int osi = (int)(q % os); q /= os;
        const int osj = (int)(q % os); q /= os;
        const int k = (int)q;
        const double w = (double)(k - wSize / 2);
        const double tj = (double)(j - cc) + (double)osj / (double)os;
        const double ti = (double)(i - cc) + (double)osi / (double)os;
        const double r2 = tj * tj + ti * ti;
        float v;
        if (w != 0.0) {
            const double fScale = sqrt(fabs(w) * wCellSize * freq0) / cellSize;
            v = (float)cos(r2 / (w * fScale));
        } else {
            v = (float)exp(-r2);
        }
        C[t] = make_float2(v, 0.f);
        acc += fabs((double)v);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(sum, acc);
}

__global__ __launch_bounds__(256) void k_scale_taps(float2 *__restrict__ C, long long n, double target,
                                                    const double *__restrict__ sum)
{
    const float sc = (float)(target / *sum);
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        C[t].x *= sc;
    }
}

}  // namespace

tcv_status launch_attic_taps(tcv_ctx *ctx, double freq0, double cellSize, int wSize, int sSize,
                             int overSample, double wCellSize, float2 *C, double *d_sum)
{
    const long long n = (long long)sSize * sSize * overSample * overSample * wSize;
    tcv_status st = launch_fill_u32(ctx, d_sum, 0, 2);
    if (st != TCV_OK) return st;
    hipLaunchKernelGGL(k_attic_taps, dim3(2048), dim3(256), 0, ctx->stream, freq0, cellSize, wSize,
                       sSize, overSample, wCellSize, C, d_sum);
    TCV_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_scale_taps, dim3(2048), dim3(256), 0, ctx->stream, C, n,
                       (double)wSize * overSample * overSample, (const double *)d_sum);
    TCV_HIP(ctx, hipGetLastError());
    return TCV_OK;
}
