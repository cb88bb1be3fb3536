// Internal declarations shared by the libtcv translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "tcv.h"

struct tcv_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t timer0 = nullptr, timer1 = nullptr;
    bool profiling = false;
    struct Pending { hipEvent_t a, b; std::string family; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    struct Stat { long launches = 0; double ms = 0.0; };
    std::map<std::string, Stat> stats;
    // device counters: [0] visibilities skipped by the plain kernels (range)
    unsigned long long *d_counters = nullptr;
    // one int per persistent workgroup: work-item hand-over without LDS
    int *d_bcast = nullptr;
    // autotuner results per workload class
    std::map<std::string, tcv_tuning> tuned;
    // major cycle: FFT plans per grid size, grow-only Hogbom workspace
    std::map<int, void *> fftPlans;
    void *majorWs = nullptr;
    size_t majorWsBytes = 0;
    // device address of the fixed scratch array of tcv_grid_noalloc on this
    // context's device (null, with the reason, when the lookup failed)
    void *scratch = nullptr;
    std::string scratchErr;
};

// Looks up the scratch array's address on the context's device (tcv_noalloc.hip).
void noalloc_setup(tcv_ctx *ctx);

// Frees the major-cycle resources of a context (tcv_major.hip).
void major_release(tcv_ctx *ctx);

// Workload class of a geometry for the autotuner cache.
std::string tuning_class(const tcv_geometry *g);

tcv_status tcv_fail(tcv_ctx *ctx, tcv_status s, const std::string &msg);

#define TCV_HIP(ctx, expr)                                                      \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess)                                                   \
            return tcv_fail((ctx), TCV_ERR_DEVICE,                              \
                            std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Brackets kernel launches with events when profiling is on.
struct ProfScope {
    tcv_ctx *ctx;
    hipEvent_t a = nullptr, b = nullptr;
    const char *family;
    ProfScope(tcv_ctx *c, const char *fam);
    ~ProfScope();
};

// Device-side view of a job's geometry (sSize copied to device memory).
struct GeomDev {
    int nVis;
    const int *iu, *iv, *cOffset, *wPlane;
    const int *sSize;  // device copy, nPlanes entries (nullptr when uniform)
    int nPlanes;
    int sUniform;      // kernel width when wPlane == nullptr
    int gSize;
    long long cSize;
};

// Launchers implemented in tcv_kernels.hip.  All enqueue on ctx->stream.
tcv_status launch_fill_u32(tcv_ctx *ctx, void *p, unsigned value, size_t n32);
tcv_status launch_copy(tcv_ctx *ctx, void *dst, const void *src, size_t n16);
tcv_status launch_read_stream(tcv_ctx *ctx, const void *buf, size_t n16, int rounds, unsigned *sink);
tcv_status launch_offsets(tcv_ctx *ctx, const double *u, const double *v,
                          const double *w, int nSamples, const double *d_freq,
                          int nChan, double cellSize, double wCellSize, int wSize,
                          int gSize, int overSample, const int *d_sSize,
                          const int *d_cOffset0, int *iu, int *iv, int *wPlane,
                          int *cOffset, unsigned long long *d_numPerPlane,
                          unsigned long long *d_bad);
tcv_status launch_grid_plain(tcv_ctx *ctx, const GeomDev &g, int maxS,
                             const float2 *data, const float2 *C, float2 *grid);
tcv_status launch_degrid_plain(tcv_ctx *ctx, const GeomDev &g,
                               const float2 *grid, const float2 *C, float2 *out);
tcv_status launch_grid_plain_f64(tcv_ctx *ctx, const GeomDev &g, const float2 *data, const float2 *C,
                                 double2 *grid);
tcv_status launch_degrid_plain_f64(tcv_ctx *ctx, const GeomDev &g, const float2 *grid, const float2 *C,
                                   double2 *out);

// ---- tiled path -------------------------------------------------------------

// One job as the persistent kernels see it (passed by value for a single
// job, or as an array in device memory for a batch).
struct JobDesc {
    const int4 *items;
    const int4 *entries;
    int *meta;                  // meta[0] = number of work items, per-XCD queues at [8..24]
    const void *data;           // visibilities to grid (float2, or half2 in FP16 storage)
    const void *C;              // taps (float2, or half2)
    float2 *grid;
    float2 *out;                // degridding output
    const float *planeScale;    // FP16 storage: per-plane tap scale
    unsigned cBytes;            // bytes of the tap table (buffer range)
    int nTX;                    // tiles per grid row
    int gSize;
    int planeTaps;              // FP16: > 0 => plane = cOffset / planeTaps
    float dataScale;            // FP16: visibility scale
    int nVis;
    // work items whose first entry's cOffset is below cEdgeLo, or whose last
    // entry's is above cEdgeHi, may hold entries on the table's first or last
    // slice (entries are ordered by C-slice key within an item)
    int cEdgeLo, cEdgeHi;
    // 1 when another job of the same launch grids into an overlapping grid: a
    // work item that holds a tile alone then still flushes with atomics
    int sharedGrid;
};

constexpr int kMaxPersistentBlocks = 8192;
constexpr size_t kBatchQueueBytes = 1024;   // ints 0 and 16 + 16 q (q < 8) are used

// Tiled kernels used when a tuning leaves the choice open.  Gridder: 1 tile in
// registers, 2 tile in LDS, 3 registers with two columns per lane.  Degridder:
// 1 registers, 2 LDS, 3 LDS with tap pairs.
constexpr int kDefaultGridKernel = 3;
constexpr int kDefaultDegridKernel = 3;
bool grid2_supported(int tw, int th);
// column-group widths of the pair gridder built for a tile shape, and the
// automatic choice (64 on 128-wide tiles, else the tile width)
bool grid_cols_supported(int tw, int th, int cols);
inline int auto_grid_cols(int tw, int th) { return tw == 128 && grid_cols_supported(tw, th, 64) ? 64 : tw; }

struct PlanShape {
    int tw, th, chunk, k2, phases, xcd;
    int kern, dkern;       // gridder / degridder kernel (codes as in tcv_tuning)
    int phasesG, xcdG;     // the gridder's work-item phases and XCD order
    int minS;              // narrowest kernel width of the job
    int tapBytes;          // stored tap size the tuning was chosen for (8 or 4)
    int itemEntries;       // entries per work item the automatic phase counts aimed at
    int gridCols;          // pair gridder: columns per load instruction (tcv_tuning.grid_cols)
    int nTX, nTY, nTiles;
    int maxS;
    long long maxEntries;  // upper bound on visibility x tile pairs
    long long nKeys;       // nTiles * k2
};

struct tcv_plan {
    PlanShape sh{};
    int nVis = 0, gSize = 0, nPlanes = 0;
    std::vector<int> sSizeHost;
    int *d_sSize = nullptr;
    int4 *entries = nullptr;      // maxEntries
    int4 *entriesTmp = nullptr;   // maxEntries: tile-grouped entries (two-level binning), or null
    int4 *items = nullptr;        // maxItems
    long long maxItems = 0;
    int *counts = nullptr;        // nKeys + 1
    int *starts = nullptr;        // nKeys + 1
    int *cursor = nullptr;        // nKeys
    int *itemCount = nullptr;     // phases * nTiles + 1
    int *itemStart = nullptr;     // phases * nTiles + 1
    int *meta = nullptr;          // 32 ints: [0] items [1] queue head [2] entries [3] bad,
                                  // [4..5] the entries' cOffset range (see note_c_range),
                                  // [8..15] per-XCD heads, [16..24] per-XCD item ranges
    // the gridder's work items (phasesG); the same buffers as above when
    // phasesG == phases
    int4 *itemsG = nullptr;
    long long maxItemsG = 0;
    int *itemCountG = nullptr, *itemStartG = nullptr, *metaG = nullptr;
    int *scanTmp = nullptr;       // scratch for the multi-level scan
    long long scanTmpLen = 0;
    long long cSize = 0;
    GeomDev geom{};               // geometry of the last build (for batch rebuilds)
    bool hasGeom = false;
    void *block = nullptr;        // one device block holding every buffer above
    bool ownsBlock = false;       // block from hipMalloc (else from a caller's arena)
};

// Plan memory as one block: plan_setup fills the shape and returns the bytes
// needed for geometry g; plan_bind carves the buffers out of such a block.
tcv_status plan_setup(tcv_ctx *ctx, const tcv_geometry *g, const tcv_tuning *tuning, tcv_plan *p,
                      size_t *bytes);
void plan_bind(tcv_plan *p, char *block);
// Whether the plan bins in two levels (by tile in LDS, then by key per tile):
// the tile histogram must fit the LDS binning kernel and the per-tile key
// histogram one workgroup's LDS.
bool plan_two_level(const PlanShape &sh);

struct tcv_batch {
    int nJobs = 0;
    int tw = 0, th = 0, kern = kDefaultGridKernel, dkern = kDefaultDegridKernel;
    int gridCols = 0;   // pair gridder column groups (PlanShape::gridCols of every plan)
    bool half = false;
    std::vector<JobDesc> host, hostD;   // gridder / degridder descriptors
    std::vector<tcv_plan *> plans;
    JobDesc *d_jobs = nullptr, *d_jobsD = nullptr;
    int *d_prefix = nullptr;      // 9 x (nJobs + 1): item prefix, then one per XCD
    int *d_queue = nullptr;       // kBatchQueueBytes: the queue and the per-XCD counters
    hipGraphExec_t graph = nullptr;
    int graphOps = 0;
};

// FP16 storage mode: taps and visibilities as half (re, im) pairs.
struct HalfStorage {
    float dataScale;            // visibility value = half * dataScale (gridding)
    const float *planeScale;    // device, per w-plane: tap value = half * planeScale[plane]
    int planeTaps;              // > 0: plane = cOffset / planeTaps; 0: plane stored in the plan
};

tcv_status plan_build_kernels(tcv_ctx *ctx, tcv_plan *p, const GeomDev &g);
// Sets JobDesc::sharedGrid on every job whose grid overlaps another job's.
void mark_shared_grids(std::vector<JobDesc> &jobs);
// half == nullptr: FP32 taps and data
JobDesc make_job_desc(const tcv_plan *p, bool forGrid, const void *data, const void *C, float2 *grid,
                      float2 *out, const HalfStorage *half);
tcv_status launch_grid_tiled(tcv_ctx *ctx, tcv_plan *p, const void *data, const void *C,
                             float2 *grid, const HalfStorage *half);
tcv_status launch_degrid_tiled(tcv_ctx *ctx, tcv_plan *p, const float2 *grid, const void *C,
                               float2 *out, const HalfStorage *half);
// one persistent launch over every job of the batch (prefix sum computed on the device)
tcv_status launch_batch(tcv_ctx *ctx, tcv_batch *b, bool degrid);
// attic C(w) taps in FP64 on the device, normalised by the FP64 sum of |C|
tcv_status launch_attic_taps(tcv_ctx *ctx, double freq0, double cellSize, int wSize, int sSize,
                             int overSample, double wCellSize, float2 *C, double *d_sum);
tcv_status launch_ctable_to_half(tcv_ctx *ctx, const float2 *C, long long nTaps, int nPlanes,
                                 const long long *d_planeStart, long long planeTaps,
                                 unsigned *d_maxbits, void *out);
tcv_status launch_vis_to_half(tcv_ctx *ctx, const float2 *data, long long n, unsigned *d_maxbits,
                              void *out);
long long scan_tmp_len(long long n);
// Supported tile configurations (tile_w, tile_h) of the tiled kernels.
bool tiled_config_supported(int tw, int th);
