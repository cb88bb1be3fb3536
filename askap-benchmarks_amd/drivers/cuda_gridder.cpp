// cuda_gridder.cpp -- replacement for attic/tConvolveCuda/CudaGridder.cc over
// libtcv: gridKernelCuda / degridKernelCuda with the signatures of
// attic/tConvolveCuda/CudaGridder.h:32-50 (std::vector arguments, a measured
// kernel time in `time`), computed on one MI355X.  The contract is the CUDA
// version's (CudaGridder.cc:54-207): every call copies its inputs to the device,
// computes, copies the result back and frees its buffers; gridding accumulates
// into `grid`, degridding overwrites `data`.  `time` covers the kernel alone
// (the device binning that precedes it, like the CUDA version's copies, is
// outside it).  Errors are reported and leave the outputs unchanged (the CUDA
// version exits the process).  Build: g++ -I include -c cuda_gridder.cpp and
// link -ltcv in place of the CUDA runtime.
#include <chrono>
#include <complex>
#include <cstdio>
#include <vector>

#include "tcv.h"

namespace {

typedef std::complex<float> Cf;

tcv_ctx *context()
{
    static tcv_ctx *c = nullptr;
    if (!c && tcv_create(0, &c) != TCV_OK) c = nullptr;
    return c;
}

// device copies of one call's arrays, freed on scope exit
struct Call {
    tcv_ctx *x;
    std::vector<void *> bufs;
    bool ok = true;
    explicit Call(tcv_ctx *c) : x(c) { ok = c != nullptr; }
    ~Call()
    {
        for (void *p : bufs) tcv_free(x, p);
    }
    void *put(const void *src, size_t bytes)
    {
        void *d = nullptr;
        if (!ok || tcv_malloc(x, bytes, &d) != TCV_OK) {
            ok = false;
            return nullptr;
        }
        bufs.push_back(d);
        if (src && bytes && tcv_upload(x, d, src, bytes) != TCV_OK) ok = false;
        return d;
    }
    void fail(const char *what)
    {
        std::fprintf(stderr, "%s: %s\n", what, x ? tcv_last_error(x) : "no device");
        ok = false;
    }
};

bool geometry(Call &c, int support, const std::vector<Cf> &C, const std::vector<int> &cOffset,
              const std::vector<int> &iu, const std::vector<int> &iv, int gSize, int *s, tcv_geometry *g)
{
    const size_t n = iu.size();
    if (iv.size() != n || cOffset.size() != n) {
        std::fprintf(stderr, "CudaGridder: iu, iv and cOffset differ in length\n");
        return false;
    }
    *s = 2 * support + 1;
    g->nVis = (int)n;
    g->d_iu = (const int *)c.put(iu.data(), n * sizeof(int));
    g->d_iv = (const int *)c.put(iv.data(), n * sizeof(int));
    g->d_cOffset = (const int *)c.put(cOffset.data(), n * sizeof(int));
    g->d_wPlane = nullptr;
    g->gSize = gSize;
    g->nPlanes = 1;
    g->sSize = s;
    g->cSize = (long long)C.size();
    return c.ok;
}

}  // namespace

void gridKernelCuda(const std::vector<Cf> &data, const int support, const std::vector<Cf> &C,
                    const std::vector<int> &cOffset, const std::vector<int> &iu,
                    const std::vector<int> &iv, std::vector<Cf> &grid, const int gSize, double &time)
{
    Call c(context());
    tcv_geometry g;
    int s = 0;
    if (!geometry(c, support, C, cOffset, iu, iv, gSize, &s, &g) || data.size() != iu.size() ||
        grid.size() != (size_t)gSize * gSize) {
        c.fail("gridKernelCuda: bad arguments");
        return;
    }
    float *dD = (float *)c.put(data.data(), data.size() * sizeof(Cf));
    float *dC = (float *)c.put(C.data(), C.size() * sizeof(Cf));
    float *dG = (float *)c.put(grid.data(), grid.size() * sizeof(Cf));
    tcv_plan *plan = nullptr;
    if (!c.ok || tcv_plan_create(c.x, &g, nullptr, &plan) != TCV_OK) return c.fail("gridKernelCuda");
    const auto t0 = std::chrono::steady_clock::now();
    const bool ok = tcv_grid_tiled(c.x, plan, dD, dC, dG) == TCV_OK && tcv_synchronize(c.x) == TCV_OK;
    time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!ok || tcv_download(c.x, grid.data(), dG, grid.size() * sizeof(Cf)) != TCV_OK) c.fail("gridKernelCuda");
    tcv_plan_destroy(c.x, plan);
}

void degridKernelCuda(const std::vector<Cf> &grid, const int gSize, const int support,
                      const std::vector<Cf> &C, const std::vector<int> &cOffset, const std::vector<int> &iu,
                      const std::vector<int> &iv, std::vector<Cf> &data, double &time)
{
    Call c(context());
    tcv_geometry g;
    int s = 0;
    if (!geometry(c, support, C, cOffset, iu, iv, gSize, &s, &g) || grid.size() != (size_t)gSize * gSize) {
        c.fail("degridKernelCuda: bad arguments");
        return;
    }
    float *dG = (float *)c.put(grid.data(), grid.size() * sizeof(Cf));
    float *dC = (float *)c.put(C.data(), C.size() * sizeof(Cf));
    float *dD = (float *)c.put(nullptr, iu.size() * sizeof(Cf));
    tcv_plan *plan = nullptr;
    if (!c.ok || tcv_plan_create(c.x, &g, nullptr, &plan) != TCV_OK) return c.fail("degridKernelCuda");
    const auto t0 = std::chrono::steady_clock::now();
    const bool ok = tcv_degrid_tiled(c.x, plan, dG, dC, dD) == TCV_OK && tcv_synchronize(c.x) == TCV_OK;
    time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<Cf> out(iu.size());
    if (ok && tcv_download(c.x, out.data(), dD, out.size() * sizeof(Cf)) == TCV_OK) data.swap(out);
    else c.fail("degridKernelCuda");
    tcv_plan_destroy(c.x, plan);
}
