/*
 * tcv.h -- C ABI of libtcv.so, the MI355X (gfx950) W-projection gridder /
 * degridder that drops in for tConvolve's Benchmark hot path.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Arguments named d_* are device pointers
 *    (allocated with tcv_malloc or any hipMalloc'd memory on the context's
 *    device); everything else is host memory.
 *  - Complex values are interleaved float pairs (re, im): the memory layout of
 *    std::complex<float> (the reference's `Value`, attic/tConvolveMPI/Benchmark.h:36).
 *  - Every call returns a tcv_status; tcv_last_error() gives the message.  The
 *    library never calls exit() or abort().
 *  - Work is queued on the context's stream and returns without waiting unless
 *    stated otherwise; tcv_synchronize() waits.  One context must not be used
 *    from two threads at once without external locking.
 *  - Gridding ACCUMULATES into the grid (as gridKernel does); degridding
 *    OVERWRITES the output visibilities (as degridKernel does).
 */
#ifndef TCV_H
#define TCV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TCV_OK = 0,
    TCV_ERR_INVALID = 1,      /* bad argument */
    TCV_ERR_RANGE = 2,        /* footprint / w-plane / C offset outside grid or table */
    TCV_ERR_DEVICE = 3,       /* HIP runtime error */
    TCV_ERR_NOMEM = 4,        /* device allocation failed */
    TCV_ERR_BUDGET = 5,       /* admission refused: HBM budget exceeded */
    TCV_ERR_UNSUPPORTED = 6,  /* shape outside what this build handles */
    TCV_ERR_CANCELLED = 7,    /* request cancelled */
    TCV_ERR_TIMEOUT = 8       /* wait timed out (the request is still pending) */
} tcv_status;

typedef struct tcv_ctx tcv_ctx;

/* ---- context ------------------------------------------------------------- */

/* Opens `device` and creates the context's stream (default priority). */
tcv_status tcv_create(int device, tcv_ctx **ctx);
void tcv_destroy(tcv_ctx *ctx);
const char *tcv_last_error(const tcv_ctx *ctx);
const char *tcv_status_string(tcv_status s);
/* Number of compute units of the context's device. */
int tcv_device_cus(const tcv_ctx *ctx);
tcv_status tcv_synchronize(tcv_ctx *ctx);

/* ---- device memory -------------------------------------------------------- */

tcv_status tcv_malloc(tcv_ctx *ctx, size_t bytes, void **d_ptr);
/* Free and total device memory in bytes. */
tcv_status tcv_mem_info(tcv_ctx *ctx, size_t *freeBytes, size_t *totalBytes);
tcv_status tcv_free(tcv_ctx *ctx, void *d_ptr);
/* Blocking copies (they synchronise the context stream first). */
tcv_status tcv_upload(tcv_ctx *ctx, void *d_dst, const void *src, size_t bytes);
tcv_status tcv_download(tcv_ctx *ctx, void *dst, const void *d_src, size_t bytes);
/* Zero fill by kernel (safe inside a captured graph). */
tcv_status tcv_zero(tcv_ctx *ctx, void *d_ptr, size_t bytes);

/* Device-to-device copy of `bytes` (a multiple of 16, both pointers 16-byte
 * aligned) by a plain streaming kernel on the context stream.  No counterpart
 * in the reference; its fixed traffic (bytes read + bytes written) calibrates
 * hardware memory-traffic counters against the gridding kernels. */
tcv_status tcv_copy(tcv_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);

/* Streaming read bandwidth of the device, in GB/s (blocking): every CU reads a
 * power-of-two buffer of `bytes` (4 KiB..64 GiB, 16-byte aligned) in 16-byte
 * loads, `rounds` (a multiple of 8) windows of 8 workgroups per CU each, timed
 * over five launches after a warm-up.  No counterpart in the reference: a
 * buffer that fits the L2s measures the L2 read bandwidth, a large one HBM's,
 * the ceilings beside which the gridding kernels' traffic is reported. */
tcv_status tcv_read_bandwidth(tcv_ctx *ctx, const void *d_buf, size_t bytes, int rounds,
                              double *gbps);

/* ---- timing (hipEvents on the context stream) ----------------------------- */

tcv_status tcv_timer_start(tcv_ctx *ctx);
/* Records the stop event, waits for it, returns the elapsed milliseconds. */
tcv_status tcv_timer_stop(tcv_ctx *ctx, double *ms);
/* When on, every gridding/degridding kernel launch is bracketed by events on
 * the launch stream; tcv_kernel_stats() waits and reports per kernel family
 * ("grid_plain", "degrid_plain", "grid_tiled", "degrid_tiled", "offsets",
 * "plan") the launch count and summed device milliseconds since the last
 * reset. */
tcv_status tcv_set_profiling(tcv_ctx *ctx, int on);
tcv_status tcv_kernel_stats(tcv_ctx *ctx, const char *family, long *launches,
                            double *total_ms);
tcv_status tcv_reset_stats(tcv_ctx *ctx);

/* ---- host set-up: the Benchmark data generators ---------------------------- */

/* Benchmark::randomInt -- attic/tConvolveMPI/Benchmark.cc:53-58 (64-bit LCG). */
int tcv_random_int(uint64_t *next);

/* Benchmark::init, u/v/w part -- attic/tConvolveMPI/Benchmark.cc:69-80. */
void tcv_attic_uvw(int nSamples, int baseline, uint64_t *next, double *u,
                   double *v, double *w);

/* Benchmark::initC (attic: real C, uniform support) --
 * attic/tConvolveMPI/Benchmark.cc:206-270.  Host computation (the table is
 * set-up, bit-identical with the reference); C == NULL only returns support
 * and wCellSize.  C holds sSize^2 * overSample^2 * wSize complex values,
 * index i + sSize*(j + sSize*(osi + os*(osj + os*k))). */
tcv_status tcv_attic_initC(double freq0, double cellSize, int wSize, int baseline,
                           int overSample, int *support, double *wCellSize,
                           float *C);

/* Problem parameters of current/tConvolveMPI runType 0..4 --
 * current/tConvolveMPI/Benchmark.cc:64-156. */
tcv_status tcv_current_params(int runType, int *gSize, double *uvCellSize,
                              int *wSize, int *support, int *overSample,
                              double *wCellSize, float *baseline);

/* ASKAP-layout uvw generator -- current/tConvolveMPI/Benchmark.cc:168-258.
 * Returns the sample count; u/v/w == NULL only sizes. */
int tcv_current_uvw(float baseline, uint64_t *next, double *u, double *v,
                    double *w);

/* Benchmark::initC (current: complex Fresnel C, per-plane sSize) --
 * current/tConvolveMPI/Benchmark.cc:471-586.  sSize/cOffset0 get wSize
 * entries (sSize[0] is an input when wSize == 1).  Returns the tap count in
 * *nTaps; C == NULL only sizes. */
tcv_status tcv_current_initC(double uvCellSize, int wSize, int support,
                             int overSample, double wCellSize, int *sSize,
                             int *cOffset0, float *C, long long *nTaps);

/* ---- a6: index set-up on the device (FP64, bit-exact) ---------------------- */

/* Benchmark::initCOffset -- attic/tConvolveMPI/Benchmark.cc:281-324 (when
 * cOffset0 == NULL: uniform kernel width sSize[0], cOffset =
 * s^2*(fracu + os*(fracv + os*woff))) and current/tConvolveMPI/Benchmark.cc:597-681
 * (cOffset0 != NULL: per-plane widths, cOffset = s_w^2*(fracu + os*fracv) +
 * cOffset0[w]; woff = 0 when wCellSize <= 0).  Visibility d = i*nChan + chan.
 * freq, sSize, cOffset0 are host arrays (nChan, wSize, wSize entries).
 * d_wPlane may be NULL; numPerPlane (host, wSize entries) may be NULL.
 * Indices are computed in FP64 without contraction, bit-identical to the
 * reference.  Blocking.  Returns TCV_ERR_RANGE (outputs still written) when
 * any w-plane is outside [0, wSize) or a footprint leaves the gSize grid; the
 * count is in *nOutOfRange when non-NULL. */
tcv_status tcv_init_offsets(tcv_ctx *ctx, const double *d_u, const double *d_v,
                            const double *d_w, int nSamples, const double *freq,
                            int nChan, double cellSize, double wCellSize,
                            int wSize, int gSize, int overSample,
                            const int *sSize, const int *cOffset0, int *d_iu,
                            int *d_iv, int *d_wPlane, int *d_cOffset,
                            long long *numPerPlane, long long *nOutOfRange);

/* ---- geometry of one gridding job ----------------------------------------- */

typedef struct {
    int nVis;                 /* visibilities (nSamples * nChan) */
    const int *d_iu;          /* centre column: columns iu-s/2 .. iu-s/2+s-1 */
    const int *d_iv;          /* first grid row of the footprint: rows iv .. iv+s-1 (the
                                 reference offsets only the column by the support) */
    const int *d_cOffset;     /* start of the visibility's C slice (complex units) */
    const int *d_wPlane;      /* w-plane per visibility; NULL => uniform width */
    int gSize;                /* grid is gSize x gSize, row-major, index iu + gSize*iv */
    int nPlanes;              /* entries in sSize (1 when uniform) */
    const int *sSize;         /* HOST: kernel width per plane (odd, 2*support+1) */
    long long cSize;          /* complex taps in C (for bounds validation) */
} tcv_geometry;

/* ---- a7/a8 plain path: the one-job-per-launch kernels ---------------------- */

/* gridKernel -- attic/tConvolveMPI/Benchmark.cc:125-155 (replaces
 * gridKernelCuda, attic/tConvolveCuda/CudaGridder.h:32-40, and
 * cuda_gridKernel, attic/tConvolveCuda/CudaGridKernel.h:33-37) with device
 * pointers: grid[iu-s/2+u, iv-s/2+v] += data * C[cOffset + v*s + u] with FP32
 * atomics, one workgroup per visibility.  Visibilities whose footprint leaves
 * the grid or whose slice leaves C are skipped and counted (see
 * tcv_range_errors). */
tcv_status tcv_grid_plain(tcv_ctx *ctx, const tcv_geometry *g,
                          const float *d_data, const float *d_C, float *d_grid);

/* degridKernel -- attic/tConvolveMPI/Benchmark.cc:158-195 (replaces
 * degridKernelCuda, attic/tConvolveCuda/CudaGridder.h:42-50): data =
 * sum grid*C over the footprint, one wave64 per visibility with a cross-lane
 * (DPP) reduction.  Skipped visibilities get 0. */
tcv_status tcv_degrid_plain(tcv_ctx *ctx, const tcv_geometry *g,
                            const float *d_grid, const float *d_C, float *d_data);

/* FP64 reference mode of the two kernels above (no counterpart in the
 * reference, which states no FP64 result; the extension every FP32 or FP16
 * result is measured against).  The same footprints over the same stored
 * complex64 taps and visibilities, promoted to double: every tap product is
 * exact and sums accumulate in FP64.  Gridding accumulates into a complex128
 * grid (interleaved double pairs) with FP64 atomics; d_data == NULL grids the
 * point-spread function (every visibility 1 + 0i).  Degridding reads a
 * complex64 grid and overwrites complex128 outputs.  Out-of-range
 * visibilities are skipped and counted as by the FP32 kernels. */
tcv_status tcv_grid_plain_f64(tcv_ctx *ctx, const tcv_geometry *g, const float *d_data,
                              const float *d_C, double *d_grid);
tcv_status tcv_degrid_plain_f64(tcv_ctx *ctx, const tcv_geometry *g, const float *d_grid,
                                const float *d_C, double *d_data);

/* Bytes of the fixed device array tcv_grid_noalloc keeps its plan in. */
#define TCV_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* tcv_grid_plain's arguments and semantics (accumulates into d_grid) computed
 * by the tiled path, with every buffer the binning plan needs (entries, work
 * items, key counters, scan partials, queue heads) in a fixed 64 MiB device
 * array that each GPU holds once for the life of the process, so the call
 * allocates no device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate -- code under stream capture, tight serving
 * loops.  The plan is the default one with its C-slice keys reduced until it
 * fits; a job whose plan still needs more than TCV_NOALLOC_SCRATCH_BYTES
 * (roughly 300 000 visibilities of width 129 on a 4096^2 grid) is refused
 * with TCV_ERR_UNSUPPORTED before anything is queued.  Every call on a device
 * uses that device's one array: calls on one context's stream run in order,
 * but calls that overlap on one GPU (two contexts, two streams) race.
 * Queued; the per-plane widths of a d_wPlane geometry are copied from the
 * host sSize array by an ordered copy on the stream. */
tcv_status tcv_grid_noalloc(tcv_ctx *ctx, const tcv_geometry *g, const float *d_data,
                            const float *d_C, float *d_grid);

/* Waits and returns (then clears) the number of visibilities skipped by the
 * kernels above because their footprint or slice was out of range. */
tcv_status tcv_range_errors(tcv_ctx *ctx, long long *count);

/* ---- a7/a8 tiled path: visibilities binned by grid tile on the device ------- */

typedef struct {
    int tile_w;      /* 8, 16, 32, 64 or 128 columns (0 = auto: 128 x 64 tiles for
                        kernels 89 wide and up when the pair kernels apply); 128-wide
                        tiles take only the pair kernels (kernel and degrid_kernel 3) */
    int tile_h;      /* rows per tile: 32 or 64 for tiles 8..64 wide, 64 for 128-wide
                        tiles, and the tall shapes 64 x 128 and 32 x 256 (pair kernels
                        only); 0 = auto */
    int chunk;       /* max entries per work item (0 = auto: 2048, 1024 for jobs with
                        kernels narrower than 24, 512 when all are) */
    int slice_keys;  /* order entries in a tile by C-slice key, 0 = no ordering */
    int phases;      /* >1: split each tile's entries into C-slice phases: equal runs of
                        the slice keys the job's entries use (0 = auto) */
    int xcd_split;   /* 1: order work items XCD-major (phases % 8 == 0), one queue per
                        XCD so each XCD's L2 holds its share of a phase's C slices;
                        0 = one queue; -1 = auto */
    int kernel;      /* gridder: 1 tile in registers, lanes = tile columns; 2 tile in
                        LDS, lanes walk each footprint's points densely; 3 registers,
                        two columns per lane (two taps per load: 16 bytes in FP32);
                        4 as 3, and four columns per lane in FP16 storage; 0 = auto */
    int degrid_kernel; /* degridder: 1 tile in registers; 2 tile in LDS, dense lanes;
                        3 LDS with tap pairs; 4 as 3, and four taps per load in FP16
                        storage; 0 = auto */
    int grid_phases; /* C-slice phases of the gridder's work items (the degridder uses
                        phases; both cut the same sorted entries); 0 = auto */
    int tap_bytes;   /* bytes per stored C tap the automatic choices are made for: 8 FP32
                        storage (0 = 8), 4 FP16 storage (tcv_grid_tiled_half and
                        tcv_degrid_tiled_half); it sets the default tile shape and phase
                        counts, which follow the table's size in bytes */
    int item_entries; /* fewest entries per work item (on average) the automatic phase
                        counts keep: finer phases keep a job's C slices in the L2s, but
                        every work item zeroes and flushes its tile (0 = 64) */
    int grid_cols;   /* pair gridder (kernel 3, FP32 and FP16 pairs): columns per load
                        instruction.  One instruction covers grid_cols columns over
                        128 / grid_cols rows; footprints narrower than the tile waste
                        fewer lanes with narrow groups.  tile_w, tile_w / 2 or tile_w / 4
                        (at least 16, and tile_h at least 512 / grid_cols); 0 = auto:
                        64 on 128-wide tiles, else tile_w */
} tcv_tuning;

typedef struct tcv_plan tcv_plan;

/* Allocates a plan for jobs of this geometry class (device buffers sized from
 * the geometry's upper bounds) and builds it from g's index arrays.  tuning
 * may be NULL: the context's tcv_autotune choice for this workload class when
 * there is one, else built-in defaults.  Zero fields of a given tuning are
 * filled with the defaults.  Blocking. */
tcv_status tcv_plan_create(tcv_ctx *ctx, const tcv_geometry *g,
                           const tcv_tuning *tuning, tcv_plan **plan);
/* Re-bins the visibilities of g (same nVis/gSize/sSize class as at creation;
 * index arrays may differ).  Queued on the stream, no host sync. */
tcv_status tcv_plan_build(tcv_ctx *ctx, tcv_plan *plan, const tcv_geometry *g);
tcv_status tcv_plan_destroy(tcv_ctx *ctx, tcv_plan *plan);
/* The tuning the plan resolved to, and after a build the number of entries
 * (visibility x tile pairs) and work items (blocking). */
tcv_status tcv_plan_info(tcv_ctx *ctx, tcv_plan *plan, tcv_tuning *tuning,
                         long long *entries, long long *items,
                         long long *outOfRange);

/* Autotuner: times candidate tunings of the tiled path on this job (hipEvents
 * on the context stream, `reps` launches each after one warm-up) and returns
 * the fastest in *best (and its average milliseconds in *bestMs).  op: 1 =
 * gridding, 2 = degridding, 3 = both (summed).  candidates == NULL runs a
 * staged search from the defaults: tile shapes suited to the kernel width x
 * the kernel families, the pair gridder's column groups, then the gridder's
 * and the degridder's C-slice phase counts from 0.6x to 4x the defaults (a job
 * that uses part of its table wants finer phases), then the work-item size.  d_grid and
 * d_out are scratch (their contents are overwritten).  The winner is cached in the context for the
 * workload class (grid size, widest kernel, table size, visibility count in
 * powers of two) and used by tcv_plan_create whenever tuning is NULL. Blocking. */
tcv_status tcv_autotune(tcv_ctx *ctx, const tcv_geometry *g, const float *d_data,
                        const float *d_C, float *d_grid, float *d_out, int op,
                        const tcv_tuning *candidates, int nCandidates, int reps,
                        tcv_tuning *best, double *bestMs);

/* Tiled gridKernel: each workgroup owns a grid tile in registers, walks the
 * tile's binned visibilities, and writes the tile back once (exclusive tiles)
 * or with FP32 atomics (tiles split over several work items). */
tcv_status tcv_grid_tiled(tcv_ctx *ctx, tcv_plan *plan, const float *d_data,
                          const float *d_C, float *d_grid);
/* Tiled degridKernel: the tile is read once into registers; each visibility's
 * partial sum over the tile is reduced with DPP and added to the output. */
tcv_status tcv_degrid_tiled(tcv_ctx *ctx, tcv_plan *plan, const float *d_grid,
                            const float *d_C, float *d_data);

/* gridKernel with isPSF = true (current/tMajorACC/tMajorACC.cc:86-126): every
 * visibility's value is 1 + 0i, so the grid accumulates the C slices alone
 * (the point-spread function's uv grid).  Tiled path; accumulates. */
tcv_status tcv_grid_psf(tcv_ctx *ctx, tcv_plan *plan, const float *d_C, float *d_grid);

/* ---- FP16 storage mode ------------------------------------------------------- */
/* C(w) taps and visibilities stored as half (re, im) pairs, scaled by the
 * max |component| of their w-plane (taps) or of the whole set (visibilities);
 * accumulation stays FP32.  No reference counterpart: the FP16 storage mode
 * is an extension of gridKernel / degridKernel that halves the C stream. */

/* Converts the complex64 table d_C (nTaps taps) into d_Chalf (4 B per tap) and
 * writes the per-plane scales to d_planeScale (nPlanes floats; host copy in
 * planeScale when non-NULL).  Planes: taps [planeStart[k], planeStart[k+1])
 * (host array, nPlanes + 1 entries), or, when planeStart is NULL, uniform
 * planes of planeTaps taps (the attic layout, planeTaps = sSize^2 * os^2).
 * Blocking. */
tcv_status tcv_ctable_to_half(tcv_ctx *ctx, const float *d_C, long long nTaps, int nPlanes,
                              const long long *planeStart, long long planeTaps,
                              void *d_Chalf, float *d_planeScale, float *planeScale);
/* Converts nVis complex64 visibilities to half pairs scaled by their max
 * |component|, returned in *scale.  Blocking. */
tcv_status tcv_vis_to_half(tcv_ctx *ctx, const float *d_data, int nVis, void *d_dhalf,
                           float *scale);
/* tcv_grid_tiled / tcv_degrid_tiled on FP16 storage.  planeTaps > 0: the
 * w-plane of a visibility is cOffset / planeTaps (uniform layout); 0: the plan
 * carries the w-plane from the geometry's d_wPlane. */
tcv_status tcv_grid_tiled_half(tcv_ctx *ctx, tcv_plan *plan, const void *d_dhalf,
                               float dataScale, const void *d_Chalf,
                               const float *d_planeScale, long long planeTaps, float *d_grid);
tcv_status tcv_degrid_tiled_half(tcv_ctx *ctx, tcv_plan *plan, const float *d_grid,
                                 const void *d_Chalf, const float *d_planeScale,
                                 long long planeTaps, float *d_data);

/* ---- batching: many jobs in one persistent launch --------------------------- */
/* No reference counterpart: the reference replicates whole processes per job
 * (attic/tConvolveMPI/tConvolveMPI.cc:78-94); here independent jobs (channels,
 * Taylor terms, snapshots) of uneven size share one launch over a device work
 * queue. */
typedef struct {
    tcv_plan *plan;        /* built plan of the job; all plans of a batch share tile_w/tile_h */
    const float *d_data;   /* complex64 visibilities to grid (may be NULL: degrid only) */
    const float *d_C;      /* complex64 C(w) taps of the job */
    float *d_grid;         /* complex64 gSize^2 grid (gridding accumulates into it) */
    float *d_out;          /* complex64 degridding output (may be NULL: grid only) */
} tcv_job;

typedef struct tcv_batch tcv_batch;

tcv_status tcv_batch_create(tcv_ctx *ctx, int nJobs, const tcv_job *jobs, tcv_batch **batch);
tcv_status tcv_batch_destroy(tcv_ctx *ctx, tcv_batch *batch);
/* Re-bins every job from the geometry its plan was last built with (queued,
 * no host sync). */
tcv_status tcv_batch_rebuild(tcv_ctx *ctx, tcv_batch *batch);
/* One persistent launch over the work items of all jobs. */
tcv_status tcv_batch_grid(tcv_ctx *ctx, tcv_batch *batch);
tcv_status tcv_batch_degrid(tcv_ctx *ctx, tcv_batch *batch);
/* ops: 1 = rebuild, 2 = grid, 4 = degrid (any sum).  The first call with a
 * given ops value captures the work into a hipGraph; later calls replay it. */
tcv_status tcv_batch_run_graph(tcv_ctx *ctx, tcv_batch *batch, int ops);

/* attic C(w) (attic/tConvolveMPI/Benchmark.cc:206-270) generated on the
 * device in FP64, normalised by the FP64 sum of |C| -- the reference sums in
 * float, which stalls near 2^24 on large tables, so these tables differ from
 * tcv_attic_initC's (bit-identical to the reference) by that normalisation
 * factor and by last-ulp libm differences.  For synthetic batch and serving
 * workloads.  Queued, no host sync. */
tcv_status tcv_attic_initC_device(tcv_ctx *ctx, double freq0, double cellSize, int wSize,
                                  int baseline, int overSample, int *support,
                                  double *wCellSize, float *d_C);

/* ---- serving: continuous batching of requests --------------------------------- */
/* No reference counterpart (the reference has no request loop; the nearest is
 * the device-resident major cycle of current/tMajorACC/tMajorACC.cc:977-1043).
 * A scheduler thread admits requests first-in first-out as the HBM arena (the
 * budget) can hold their buffers, batches every admitted request into one
 * persistent launch on a free stream slot, and overlaps up to maxInflight
 * batches on the device.  C tables, visibilities and grids to degrid stay
 * resident with the caller; the arena holds each request's plan and its result
 * (a new grid, or degridded visibilities). */
typedef struct tcv_server tcv_server;

typedef struct {
    size_t arenaBytes;   /* HBM budget for request buffers, allocated once */
    int maxInflight;     /* batches concurrently on the device (1..8 streams) */
    int maxBatch;        /* requests per batch (0 = all admitted) */
    int tile_w, tile_h, chunk, slice_keys, phases;   /* plan tuning (0 = auto) */
    int autoRelease;     /* 1: free each result when its request completes */
    long long maxBatchPoints; /* a batch closes once it holds this many grid points, so
                                 large requests run apart from small ones (0 = no limit) */
    int item_entries;    /* plan tuning: tcv_tuning.item_entries (0 = auto) */
    int grid_cols;       /* plan tuning: tcv_tuning.grid_cols (0 = auto) */
    long long exclusivePoints; /* a request of at least this many grid points runs alone: it
                                  waits for the running batches to finish, and no batch starts
                                  beside it (0 = never) */
} tcv_server_config;

typedef struct {
    int op;                /* 0 = grid into a new zeroed grid, 1 = degrid d_grid */
    tcv_geometry geom;     /* resident index arrays (geom.sSize is copied) */
    const float *d_data;   /* op 0: complex64 visibilities */
    const float *d_C;      /* complex64 C(w) taps */
    const float *d_grid;   /* op 1: complex64 gSize^2 grid */
} tcv_request;

typedef struct {
    double submitted, admitted, launched, completed;   /* seconds, server clock */
    tcv_status status;     /* TCV_OK, TCV_ERR_CANCELLED or the failure */
    const float *d_result; /* op 0: gSize^2 grid; op 1: nVis visibilities (NULL after release) */
    long long points;      /* nVis * sSize^2 (largest width when widths vary) */
} tcv_result;

tcv_status tcv_server_create(tcv_ctx *ctx, const tcv_server_config *cfg, tcv_server **srv);
/* Stops the scheduler after the batches on the device finish; frees the arena. */
tcv_status tcv_server_destroy(tcv_server *srv);
/* Arena bytes a request would hold while it runs (its plan plus its result),
 * the amount admission waits for.  Validates the request; enqueues nothing. */
tcv_status tcv_server_footprint(tcv_server *srv, const tcv_request *req, size_t *bytes);
/* Thread-safe.  TCV_ERR_BUDGET when the request can never fit the budget. */
tcv_status tcv_server_submit(tcv_server *srv, const tcv_request *req, long long *id);
/* Cancels a request that is still queued (TCV_ERR_UNSUPPORTED once admitted). */
tcv_status tcv_server_cancel(tcv_server *srv, long long id);
/* Waits until the request completed, failed or was cancelled, at most
 * timeout_s seconds (TCV_ERR_TIMEOUT, request still pending). */
tcv_status tcv_server_wait(tcv_server *srv, long long id, double timeout_s, tcv_result *res);
/* Frees a finished request's result buffer and forgets the request. */
tcv_status tcv_server_release(tcv_server *srv, long long id);
tcv_status tcv_server_stats(tcv_server *srv, long long *queued, long long *running,
                            size_t *arenaUsed, size_t *arenaCapacity);
double tcv_server_clock(tcv_server *srv);
const char *tcv_server_last_error(tcv_server *srv);


/* ---- major cycle (current/tMajorACC/tMajorACC.cc) -------------------------
 * Device-resident neighbours of the gridder in an imaging major cycle.  All
 * calls are queued on the context stream unless marked blocking. */

/* fftExec (tMajorACC.cc:488-546, GPU form :548-618): in-place unnormalised
 * 2-D C2C transform of a gSize x gSize complex grid with the origin at pixel
 * (gSize/2, gSize/2) (the reference rotates by gSize/2 before and after);
 * forward != 0 uses e^{-i} (FFTW_FORWARD), else e^{+i}.  gSize must be even
 * (the reference refuses odd sizes too).  hipFFT plans are cached per size. */
tcv_status tcv_fft_exec(tcv_ctx *ctx, float *d_grid, int gSize, int forward);
/* fftFix (tMajorACC.cc:620-636): grid = real(grid) * scale, imaginary part 0. */
tcv_status tcv_fft_fix(tcv_ctx *ctx, float *d_grid, long long nPixels, float scale);
/* tcv_fft_exec followed by tcv_fft_fix, the fix folded into the transform's
 * last pass (the reference's inverse-FFT-then-fftFix pair, tMajorACC.cc:1121-1140). */
tcv_status tcv_fft_exec_fix(tcv_ctx *ctx, float *d_grid, int gSize, int forward, float scale);
/* findPeak (tMajorACC.cc:689-701): the first pixel of largest |real part| of
 * a complex image; *maxVal is its (signed) real part.  Blocking. */
tcv_status tcv_find_peak(tcv_ctx *ctx, const float *d_image, long long nPixels,
                         float *maxVal, long long *maxPos);
/* deconvolve (Hogbom CLEAN, tMajorACC.cc:800-840, subtractPsf :723-750):
 * niters iterations on the complex residual (width x width) with the complex
 * PSF image (psfWidth x psfWidth), adding gain * peak to the complex model.
 * Only real parts change, rounded as the reference rounds them; the PSF row
 * offset follows the reference (ry - px), and PSF rows it would place outside
 * the PSF are skipped.  No host round trip per iteration.  The four peak
 * outputs (PSF peak, first residual peak) may be NULL; when any is given the
 * call blocks until the loop has finished.  Widths below 65536. */
tcv_status tcv_hogbom(tcv_ctx *ctx, float *d_residual, long long width, const float *d_psf,
                      long long psfWidth, float *d_model, int niters, float gain,
                      float *psfPeakVal, long long *psfPeakPos, float *dirtyPeakVal,
                      long long *dirtyPeakPos);
/* data = data - model for nVis complex visibilities (tMajorACC.cc:1549-1556). */
tcv_status tcv_vis_subtract(tcv_ctx *ctx, float *d_data, const float *d_model, long long nVis);

#ifdef __cplusplus
}
#endif
#endif
