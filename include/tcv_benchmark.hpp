// tcv_benchmark.hpp -- header-only C++ drop-in for tConvolve's Benchmark class,
// running the hot path on one MI355X through the C ABI in tcv.h.
//
// tcv::Benchmark has the public surface of attic/tConvolveMPI/Benchmark.h:55-82
// (randomInt, init, runGrid, runDegrid, gridKernel, degridKernel, initC,
// initCOffset, getSupport); tcv::CurrentBenchmark that of
// current/tConvolveMPI/Benchmark.h:38-72 (adding nVisibilitiesGridded,
// nPixelsGridded, requiredRate, setMPIrank, setSort, setRunType, getRunType).
// include/tconvolve/{attic,current}/Benchmark.h put them in the global
// namespace under the reference's header name, so the reference's drivers
// compile against them unchanged.
// The compile-time constants of attic/tConvolveMPI/Benchmark.h:39-46 become
// the fields of tcv::Config (defaults = the reference's values).
//
// Residency: init() uploads C(w) and computes the offsets on the device;
// runGrid()/runDegrid() work on device-resident state and wait for completion
// (the reference's calls are synchronous).  The explicit gridKernel /
// degridKernel overloads taking std::vectors keep the reference semantics
// exactly: they copy in, compute on the device, and copy the result back.
#ifndef TCV_BENCHMARK_HPP
#define TCV_BENCHMARK_HPP

#include <complex>
#include <cstdint>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "tcv.h"

namespace tcv {

typedef double Coord;
typedef float Real;
typedef std::complex<Real> Value;

class Error : public std::runtime_error {
public:
    Error(tcv_status s, const std::string &m) : std::runtime_error(m), status(s) {}
    tcv_status status;
};

inline void check(tcv_ctx *ctx, tcv_status s, const char *what)
{
    if (s != TCV_OK)
        throw Error(s, std::string(what) + ": " + tcv_status_string(s) + " (" +
                           (ctx ? tcv_last_error(ctx) : "") + ")");
}

// RAII device buffer of n elements of T.
template <class T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(tcv_ctx *ctx, size_t n) : ctx_(ctx), n_(n)
    {
        void *p = nullptr;
        check(ctx, tcv_malloc(ctx, n * sizeof(T), &p), "tcv_malloc");
        p_ = static_cast<T *>(p);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept { swap(o); }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
    {
        reset();
        swap(o);
        return *this;
    }
    ~DeviceBuffer() { reset(); }
    void reset()
    {
        if (p_) tcv_free(ctx_, p_);
        p_ = nullptr;
        n_ = 0;
    }
    void upload(const T *src, size_t n)
    {
        if (n > n_) throw Error(TCV_ERR_INVALID, "upload larger than buffer");
        check(ctx_, tcv_upload(ctx_, p_, src, n * sizeof(T)), "tcv_upload");
    }
    void download(T *dst, size_t n) const
    {
        if (n > n_) throw Error(TCV_ERR_INVALID, "download larger than buffer");
        check(ctx_, tcv_download(ctx_, dst, p_, n * sizeof(T)), "tcv_download");
    }
    void zero() { check(ctx_, tcv_zero(ctx_, p_, n_ * sizeof(T)), "tcv_zero"); }
    T *get() const { return p_; }
    size_t size() const { return n_; }

private:
    void swap(DeviceBuffer &o)
    {
        std::swap(ctx_, o.ctx_);
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }
    tcv_ctx *ctx_ = nullptr;
    T *p_ = nullptr;
    size_t n_ = 0;
};

// attic/tConvolveMPI/Benchmark.h:38-46
struct Config {
    int nSamples = 160000;   // Number of data samples
    int wSize = 33;          // Number of lookup planes in w projection
    int nChan = 1;           // Number of spectral channels
    int gSize = 4096;        // Size of output grid in pixels
    Coord cellSize = 5.0;    // Cellsize of output grid in wavelengths
    int baseline = 2000;     // Maximum baseline in meters
    bool tiled = true;       // tiled (binned) kernels; false = plain atomic kernels
    int device = 0;
};

class Benchmark {
public:
    explicit Benchmark(const Config &cfg = Config()) : cfg_(cfg), next(1)
    {
        check(nullptr, tcv_create(cfg.device, &ctx_), "tcv_create");
    }
    ~Benchmark()
    {
        if (plan_) tcv_plan_destroy(ctx_, plan_);
        grid_.reset(); C_.reset(); data_.reset(); outdata_.reset();
        iu_.reset(); iv_.reset(); cOffset_.reset();
        tcv_destroy(ctx_);
    }
    Benchmark(const Benchmark &) = delete;
    Benchmark &operator=(const Benchmark &) = delete;

    // attic/tConvolveMPI/Benchmark.cc:53-58
    int randomInt() { return tcv_random_int(&next); }

    // attic/tConvolveMPI/Benchmark.cc:60-96
    void init()
    {
        const int n = cfg_.nSamples, nv = cfg_.nSamples * cfg_.nChan;
        u.resize(n);
        v.resize(n);
        w.resize(n);
        const Coord maxint = std::numeric_limits<int>::max();
        for (int i = 0; i < n; i++) {
            u[i] = cfg_.baseline * Coord(randomInt()) / maxint - cfg_.baseline / 2;
            v[i] = cfg_.baseline * Coord(randomInt()) / maxint - cfg_.baseline / 2;
            w[i] = cfg_.baseline * Coord(randomInt()) / maxint - cfg_.baseline / 2;
        }
        std::vector<Value> ones(nv, Value(1.0f, 0.0f));
        data_ = DeviceBuffer<Value>(ctx_, nv);
        data_.upload(ones.data(), nv);
        outdata_ = DeviceBuffer<Value>(ctx_, nv);
        outdata_.zero();
        grid_ = DeviceBuffer<Value>(ctx_, size_t(cfg_.gSize) * cfg_.gSize);
        grid_.zero();
        std::vector<Coord> freq(cfg_.nChan);
        for (int i = 0; i < cfg_.nChan; i++)
            freq[i] = (1.4e9 - 2.0e5 * Coord(i) / Coord(cfg_.nChan)) / 2.998e8;
        std::vector<Value> C;
        initC(freq, cfg_.cellSize, cfg_.wSize, m_support, overSample, wCellSize, C);
        C_ = DeviceBuffer<Value>(ctx_, C.size());
        C_.upload(C.data(), C.size());
        initCOffset(u, v, w, freq, cfg_.cellSize, wCellSize, cfg_.wSize, cfg_.gSize, m_support,
                    overSample);
    }

    void runGrid()
    {
        gridResident();
        check(ctx_, tcv_synchronize(ctx_), "runGrid");
    }

    void runDegrid()
    {
        degridResident();
        check(ctx_, tcv_synchronize(ctx_), "runDegrid");
    }

    // attic/tConvolveMPI/Benchmark.cc:125-155: accumulate data*C into a host grid
    // (the samples are the ones initCOffset produced).
    void gridKernel(const int support, const std::vector<Value> &C, std::vector<Value> &grid,
                    const int gSize)
    {
        requireGeometry(support, gSize);
        DeviceBuffer<Value> dC(ctx_, C.size()), dG(ctx_, grid.size());
        dC.upload(C.data(), C.size());
        dG.upload(grid.data(), grid.size());
        launchGrid(dC.get(), dG.get(), (long long)C.size());
        dG.download(grid.data(), grid.size());
    }

    // attic/tConvolveMPI/Benchmark.cc:158-195: data = sum grid*C (overwrites data)
    void degridKernel(const std::vector<Value> &grid, const int gSize, const int support,
                      const std::vector<Value> &C, std::vector<Value> &data)
    {
        requireGeometry(support, gSize);
        DeviceBuffer<Value> dC(ctx_, C.size()), dG(ctx_, grid.size()), dD(ctx_, nVis());
        dC.upload(C.data(), C.size());
        dG.upload(grid.data(), grid.size());
        launchDegrid(dG.get(), dC.get(), dD.get(), (long long)C.size());
        data.resize(nVis());
        dD.download(data.data(), data.size());
    }

    // attic/tConvolveMPI/Benchmark.cc:206-270 (host, bit-identical to the reference)
    void initC(const std::vector<Coord> &freq, const Coord cellSize, const int wSize,
               int &support, int &overSample_, Coord &wCellSize_, std::vector<Value> &C)
    {
        overSample_ = 8;
        check(ctx_, tcv_attic_initC(freq[0], cellSize, wSize, cfg_.baseline, overSample_, &support,
                                    &wCellSize_, nullptr), "tcv_attic_initC");
        const size_t s = 2 * support + 1;
        C.resize(s * s * overSample_ * overSample_ * wSize);
        check(ctx_, tcv_attic_initC(freq[0], cellSize, wSize, cfg_.baseline, overSample_,
                                    &support, &wCellSize_, reinterpret_cast<float *>(C.data())),
              "tcv_attic_initC");
    }

    // attic/tConvolveMPI/Benchmark.cc:281-324, on the device in FP64
    void initCOffset(const std::vector<Coord> &u_, const std::vector<Coord> &v_,
                     const std::vector<Coord> &w_, const std::vector<Coord> &freq,
                     const Coord cellSize, const Coord wCellSize_, const int wSize,
                     const int gSize, const int support, const int overSample_)
    {
        const int ns = (int)u_.size(), nc = (int)freq.size(), nv = ns * nc;
        DeviceBuffer<Coord> du(ctx_, ns), dv(ctx_, ns), dw(ctx_, ns);
        du.upload(u_.data(), ns);
        dv.upload(v_.data(), ns);
        dw.upload(w_.data(), ns);
        iu_ = DeviceBuffer<int>(ctx_, nv);
        iv_ = DeviceBuffer<int>(ctx_, nv);
        cOffset_ = DeviceBuffer<int>(ctx_, nv);
        sSize_ = 2 * support + 1;
        gSize_ = gSize;
        nVis_ = nv;
        cTaps_ = (long long)sSize_ * sSize_ * overSample_ * overSample_ * wSize;
        long long bad = 0;
        tcv_status st = tcv_init_offsets(ctx_, du.get(), dv.get(), dw.get(), ns, freq.data(), nc,
                                         cellSize, wCellSize_, wSize, gSize, overSample_, &sSize_,
                                         nullptr, iu_.get(), iv_.get(), nullptr, cOffset_.get(),
                                         nullptr, &bad);
        // out-of-range visibilities are counted here and skipped by the kernels
        if (st != TCV_OK && st != TCV_ERR_RANGE) check(ctx_, st, "tcv_init_offsets");
        nOutOfRange = bad;
        if (plan_) {
            tcv_plan_destroy(ctx_, plan_);
            plan_ = nullptr;
        }
    }

    int getSupport() { return m_support; }

    // ---- extras beyond the reference surface ------------------------------
    std::vector<Value> grid() const
    {
        std::vector<Value> g(grid_.size());
        grid_.download(g.data(), g.size());
        return g;
    }
    std::vector<Value> outdata() const
    {
        std::vector<Value> d(outdata_.size());
        outdata_.download(d.data(), d.size());
        return d;
    }
    long long gridPoints() const { return (long long)nVis_ * sSize_ * sSize_; }
    tcv_ctx *context() const { return ctx_; }
    long long nOutOfRange = 0;

private:
    int nVis() const { return nVis_; }
    void requireGeometry(int support, int gSize) const
    {
        if (2 * support + 1 != sSize_ || gSize != gSize_)
            throw Error(TCV_ERR_INVALID, "support/gSize differ from the initialised problem");
    }
    tcv_geometry geometry(long long cTaps) const
    {
        tcv_geometry g;
        g.nVis = nVis_;
        g.d_iu = iu_.get();
        g.d_iv = iv_.get();
        g.d_cOffset = cOffset_.get();
        g.d_wPlane = nullptr;
        g.gSize = gSize_;
        g.nPlanes = 1;
        g.sSize = &sSize_;
        g.cSize = cTaps;
        return g;
    }
    tcv_plan *plan(long long cTaps)
    {
        if (!plan_) {
            tcv_geometry g = geometry(cTaps);
            check(ctx_, tcv_plan_create(ctx_, &g, nullptr, &plan_), "tcv_plan_create");
        }
        return plan_;
    }
    void launchGrid(Value *dC, Value *dG, long long cTaps)
    {
        if (cfg_.tiled) {
            check(ctx_, tcv_grid_tiled(ctx_, plan(cTaps), reinterpret_cast<float *>(data_.get()),
                                       reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dG)),
                  "tcv_grid_tiled");
        } else {
            tcv_geometry g = geometry(cTaps);
            check(ctx_, tcv_grid_plain(ctx_, &g, reinterpret_cast<float *>(data_.get()),
                                       reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dG)),
                  "tcv_grid_plain");
        }
    }
    void launchDegrid(Value *dG, Value *dC, Value *dD, long long cTaps)
    {
        if (cfg_.tiled) {
            check(ctx_, tcv_degrid_tiled(ctx_, plan(cTaps), reinterpret_cast<float *>(dG),
                                         reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dD)),
                  "tcv_degrid_tiled");
        } else {
            tcv_geometry g = geometry(cTaps);
            check(ctx_, tcv_degrid_plain(ctx_, &g, reinterpret_cast<float *>(dG),
                                         reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dD)),
                  "tcv_degrid_plain");
        }
    }
    void gridResident() { launchGrid(C_.get(), grid_.get(), (long long)C_.size()); }
    void degridResident() { launchDegrid(grid_.get(), C_.get(), outdata_.get(), (long long)C_.size()); }

    Config cfg_;
    tcv_ctx *ctx_ = nullptr;
    tcv_plan *plan_ = nullptr;
    DeviceBuffer<Value> grid_, C_, data_, outdata_;
    DeviceBuffer<int> iu_, iv_, cOffset_;
    int sSize_ = 0, gSize_ = 0, nVis_ = 0;
    long long cTaps_ = 0;

public:
    // the reference's members (attic/tConvolveMPI/Benchmark.h:84-98)
    std::vector<Coord> u, v, w;
    int m_support = 0;
    int overSample = 0;
    Coord wCellSize = 0.0;
    uint64_t next;   // LCG state, seed 1
};

// Options of the current-generation drop-in (the reference picks its problem by
// setRunType; these only choose the device and the kernels).
struct CurrentConfig {
    bool tiled = true;   // tiled (binned) kernels; false = plain atomic kernels
    int device = 0;
};

// current/tConvolveMPI/Benchmark.h:38-72: per-plane kernel widths with complex
// Fresnel C(w), the ASKAP-layout generator, runType 0..4, setSort, the
// gridded-pixel count and the ASKAP required rates.  State lives on the device
// (C, the index arrays, grid1, outdata1); runGrid()/runDegrid() wait for the
// device, as the reference's calls return when done.
class CurrentBenchmark {
public:
    explicit CurrentBenchmark(const CurrentConfig &cfg = CurrentConfig()) : cfg_(cfg), next(1)
    {
        check(nullptr, tcv_create(cfg.device, &ctx_), "tcv_create");
    }
    ~CurrentBenchmark()
    {
        if (plan_) tcv_plan_destroy(ctx_, plan_);
        grid1_.reset(); C_.reset(); data_.reset(); outdata1_.reset();
        iu_.reset(); iv_.reset(); wPlane_.reset(); cOffset_.reset();
        tcv_destroy(ctx_);
    }
    CurrentBenchmark(const CurrentBenchmark &) = delete;
    CurrentBenchmark &operator=(const CurrentBenchmark &) = delete;

    // current/tConvolveMPI/Benchmark.cc:54-59
    int randomInt() { return tcv_random_int(&next); }

    // current/tConvolveMPI/Benchmark.cc:61-300: problem parameters of the run type,
    // the ASKAP-layout u, v, w, C(w) with per-plane widths, the offsets (device,
    // FP64) and, with setSort(1), the index arrays ordered by w-plane
    void init()
    {
        float bl = 0.f;
        check(ctx_, tcv_current_params(runType, &gSize, &uvCellSize, &wSize, &m_support, &overSample,
                                       &wCellSize, &bl), "tcv_current_params");
        baseline = bl;
        uint64_t probe = next;   // sizing only: the generator's state moves once, below
        nSamples = tcv_current_uvw(baseline, &probe, nullptr, nullptr, nullptr);
        u.resize(nSamples);
        v.resize(nSamples);
        w.resize(nSamples);
        tcv_current_uvw(baseline, &next, u.data(), v.data(), w.data());
        nChan = 1;
        const int nv = nSamples * nChan;
        std::vector<Value> ones(nv, Value(1.0f, 0.0f));
        data_ = DeviceBuffer<Value>(ctx_, nv);
        data_.upload(ones.data(), nv);
        outdata1_ = DeviceBuffer<Value>(ctx_, nv);
        outdata1_.zero();
        grid1_ = DeviceBuffer<Value>(ctx_, size_t(gSize) * gSize);
        grid1_.zero();
        // current/tConvolveMPI/Benchmark.cc:266-269: maxFreqHz is a float, 1 GHz
        std::vector<Coord> wavenumber(nChan);
        for (int i = 0; i < nChan; i++)
            wavenumber[i] = (Coord(float(1.0e9)) - 2.0e5 * Coord(i) / Coord(nChan)) / 2.998e8;
        std::vector<Value> C;
        initC(uvCellSize, wSize, m_support, overSample, wCellSize, C);
        C_ = DeviceBuffer<Value>(ctx_, C.size());
        C_.upload(C.data(), C.size());
        cTaps_ = (long long)C.size();
        initCOffset(u, v, w, wavenumber, uvCellSize, wCellSize, wSize, gSize, overSample);
        if (doSort == 1 && wSize > 1) sortByPlane();
    }

    void runGrid()
    {
        launchGrid(C_.get(), grid1_.get());
        check(ctx_, tcv_synchronize(ctx_), "runGrid");
    }
    void runDegrid()
    {
        launchDegrid(grid1_.get(), C_.get(), outdata1_.get());
        check(ctx_, tcv_synchronize(ctx_), "runDegrid");
    }

    // current/tConvolveMPI/Benchmark.cc:373-413 on host vectors: the samples are
    // the ones init produced; data = 1 + 0i; accumulates into grid
    void gridKernel(const std::vector<Value> &C, std::vector<Value> &grid, const int gSize_)
    {
        requireGeometry(C, gSize_);
        DeviceBuffer<Value> dC(ctx_, C.size()), dG(ctx_, grid.size());
        dC.upload(C.data(), C.size());
        dG.upload(grid.data(), grid.size());
        launchGrid(dC.get(), dG.get());
        dG.download(grid.data(), grid.size());
    }
    // current/tConvolveMPI/Benchmark.cc:416-460: data = sum grid * C (overwrites)
    void degridKernel(const std::vector<Value> &grid, const int gSize_, const std::vector<Value> &C,
                      std::vector<Value> &data)
    {
        requireGeometry(C, gSize_);
        const int nv = nSamples * nChan;
        DeviceBuffer<Value> dC(ctx_, C.size()), dG(ctx_, grid.size()), dD(ctx_, nv);
        dC.upload(C.data(), C.size());
        dG.upload(grid.data(), grid.size());
        launchDegrid(dG.get(), dC.get(), dD.get());
        data.resize(nv);
        dD.download(data.data(), data.size());
    }

    // current/tConvolveMPI/Benchmark.cc:471-586 (host, bit-identical to the
    // reference): fills the per-plane widths and plane offsets and C
    void initC(const Coord uvCellSize_, const int wSize_, int &support, int &overSample_,
               Coord &wCellSize_, std::vector<Value> &C)
    {
        sSize.assign(wSize_, 2 * support + 1);
        cOffset0.assign(wSize_, 0);
        long long n = 0;
        check(ctx_, tcv_current_initC(uvCellSize_, wSize_, support, overSample_, wCellSize_, sSize.data(),
                                      cOffset0.data(), nullptr, &n), "tcv_current_initC");
        C.resize(size_t(n));
        check(ctx_, tcv_current_initC(uvCellSize_, wSize_, support, overSample_, wCellSize_, sSize.data(),
                                      cOffset0.data(), reinterpret_cast<float *>(C.data()), &n),
              "tcv_current_initC");
    }

    // current/tConvolveMPI/Benchmark.cc:597-681, on the device in FP64
    void initCOffset(const std::vector<Coord> &u_, const std::vector<Coord> &v_,
                     const std::vector<Coord> &w_, const std::vector<Coord> &freq,
                     const Coord uvCellSize_, const Coord wCellSize_, const int wSize_,
                     const int gSize_, const int overSample_)
    {
        const int ns = (int)u_.size(), nc = (int)freq.size(), nv = ns * nc;
        DeviceBuffer<Coord> du(ctx_, ns), dv(ctx_, ns), dw(ctx_, ns);
        du.upload(u_.data(), ns);
        dv.upload(v_.data(), ns);
        dw.upload(w_.data(), ns);
        iu_ = DeviceBuffer<int>(ctx_, nv);
        iv_ = DeviceBuffer<int>(ctx_, nv);
        wPlane_ = DeviceBuffer<int>(ctx_, nv);
        cOffset_ = DeviceBuffer<int>(ctx_, nv);
        numPerPlane.assign(wSize_, 0);
        long long bad = 0;
        tcv_status st = tcv_init_offsets(ctx_, du.get(), dv.get(), dw.get(), ns, freq.data(), nc, uvCellSize_,
                                         wCellSize_, wSize_, gSize_, overSample_, sSize.data(),
                                         cOffset0.data(), iu_.get(), iv_.get(), wPlane_.get(),
                                         cOffset_.get(), numPerPlane.data(), &bad);
        // out-of-range visibilities are counted here and skipped by the kernels
        if (st != TCV_OK && st != TCV_ERR_RANGE) check(ctx_, st, "tcv_init_offsets");
        nOutOfRange = bad;
        if (plan_) {
            tcv_plan_destroy(ctx_, plan_);
            plan_ = nullptr;
        }
    }

    int getSupport() { return m_support; }
    long nVisibilitiesGridded() { return long(nSamples) * nChan; }
    // current/tConvolveMPI/Benchmark.cc:683-700
    long nPixelsGridded()
    {
        long vis = 0, pix = 0;
        for (int k = 0; k < wSize; k++) {
            vis += long(numPerPlane[k]);
            pix += long(numPerPlane[k]) * long(sSize[k]) * long(sSize[k]);
        }
        if (vis != nVisibilitiesGridded()) {
            std::cerr << "Visibility count error: " << vis << " != " << nVisibilitiesGridded() << std::endl;
            return -1;
        }
        return pix;
    }
    // current/tConvolveMPI/Benchmark.cc:702-740: the ASKAP gridding rates a
    // process must sustain (continuum, spectral line), in pixels per second
    std::vector<float> requiredRate()
    {
        std::vector<float> rates(2);
        double tmax = 5250.;
        long Nvis = (36 * 35) / 2 * 12 * 3600. / 5.;
        long Ncycles = 10, Npercycle = 3, NTT = 3, Npol = 1, Nchanperproc = 1, Npixels = 20 * 20;
        const long efficiency = 10;
        rates[0] = float(efficiency * Nvis * Ncycles * Npercycle * NTT * Npol * Nchanperproc * Npixels / tmax);
        std::cout << "continuum gridding rate (per process) for "
                  << Nvis * Ncycles * Npercycle * NTT * Npol * Nchanperproc * Npixels / 1e9 << "e9 pix gridded is "
                  << rates[0] / 1e6 << " Mpix/sec" << std::endl;
        tmax = 5500.;
        Nvis = (30 * 29) / 2 * 12 * 3600. / 5.;
        Ncycles = 4;
        NTT = 1;
        Nchanperproc = 78;
        Npixels = 10 * 10;
        rates[1] = float(efficiency * Nvis * Ncycles * Npercycle * NTT * Npol * Nchanperproc * Npixels / tmax);
        std::cout << "spectral gridding rate (per process) for "
                  << Nvis * Ncycles * Npercycle * NTT * Npol * Nchanperproc * Npixels / 1e9 << "e9 pix gridded is "
                  << rates[1] / 1e6 << " Mpix/sec" << std::endl;
        return rates;
    }

    void setMPIrank(const int rank) { mpirank = rank; }
    void setSort(const int type) { doSort = type; }
    void setRunType(const int type) { runType = type; }
    int getRunType() { return runType; }

    // ---- extras beyond the reference surface ------------------------------
    std::vector<Value> grid() const
    {
        std::vector<Value> g(grid1_.size());
        grid1_.download(g.data(), g.size());
        return g;
    }
    std::vector<Value> outdata() const
    {
        std::vector<Value> d(outdata1_.size());
        outdata1_.download(d.data(), d.size());
        return d;
    }
    // the device index arrays (after any sort), for checks
    void samples(std::vector<int> &iu, std::vector<int> &iv, std::vector<int> &wPlane,
                 std::vector<int> &cOffset) const
    {
        const size_t nv = iu_.size();
        iu.resize(nv); iv.resize(nv); wPlane.resize(nv); cOffset.resize(nv);
        iu_.download(iu.data(), nv);
        iv_.download(iv.data(), nv);
        wPlane_.download(wPlane.data(), nv);
        cOffset_.download(cOffset.data(), nv);
    }
    tcv_ctx *context() const { return ctx_; }
    long long nOutOfRange = 0;

private:
    // current/tConvolveMPI/Benchmark.cc:275-298: stable counting sort by w-plane
    // of iu, iv, wPlane, cOffset (the data, all ones, is not permuted)
    void sortByPlane()
    {
        std::vector<int> iu, iv, wp, co;
        samples(iu, iv, wp, co);
        std::vector<long long> at(wSize, 0);
        for (int k = 1; k < wSize; k++) at[k] = at[k - 1] + numPerPlane[k - 1];
        std::vector<int> iu2(iu.size()), iv2(iu.size()), wp2(iu.size()), co2(iu.size());
        for (size_t i = 0; i < iu.size(); i++) {
            const long long j = at[wp[i]]++;
            iu2[j] = iu[i];
            iv2[j] = iv[i];
            wp2[j] = wp[i];
            co2[j] = co[i];
        }
        iu_.upload(iu2.data(), iu2.size());
        iv_.upload(iv2.data(), iv2.size());
        wPlane_.upload(wp2.data(), wp2.size());
        cOffset_.upload(co2.data(), co2.size());
    }
    void requireGeometry(const std::vector<Value> &C, int gSize_) const
    {
        if (gSize_ != gSize || (long long)C.size() != cTaps_)
            throw Error(TCV_ERR_INVALID, "C/gSize differ from the initialised problem");
    }
    tcv_geometry geometry() const
    {
        tcv_geometry g;
        g.nVis = nSamples * nChan;
        g.d_iu = iu_.get();
        g.d_iv = iv_.get();
        g.d_cOffset = cOffset_.get();
        g.d_wPlane = wPlane_.get();
        g.gSize = gSize;
        g.nPlanes = wSize;
        g.sSize = sSize.data();
        g.cSize = cTaps_;
        return g;
    }
    tcv_plan *plan()
    {
        if (!plan_) {
            tcv_geometry g = geometry();
            check(ctx_, tcv_plan_create(ctx_, &g, nullptr, &plan_), "tcv_plan_create");
        }
        return plan_;
    }
    void launchGrid(Value *dC, Value *dG)
    {
        if (cfg_.tiled) {
            check(ctx_, tcv_grid_tiled(ctx_, plan(), reinterpret_cast<float *>(data_.get()),
                                       reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dG)),
                  "tcv_grid_tiled");
        } else {
            tcv_geometry g = geometry();
            check(ctx_, tcv_grid_plain(ctx_, &g, reinterpret_cast<float *>(data_.get()),
                                       reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dG)),
                  "tcv_grid_plain");
        }
    }
    void launchDegrid(Value *dG, Value *dC, Value *dD)
    {
        if (cfg_.tiled) {
            check(ctx_, tcv_degrid_tiled(ctx_, plan(), reinterpret_cast<float *>(dG),
                                         reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dD)),
                  "tcv_degrid_tiled");
        } else {
            tcv_geometry g = geometry();
            check(ctx_, tcv_degrid_plain(ctx_, &g, reinterpret_cast<float *>(dG),
                                         reinterpret_cast<float *>(dC), reinterpret_cast<float *>(dD)),
                  "tcv_degrid_plain");
        }
    }

    CurrentConfig cfg_;
    tcv_ctx *ctx_ = nullptr;
    tcv_plan *plan_ = nullptr;
    DeviceBuffer<Value> grid1_, C_, data_, outdata1_;
    DeviceBuffer<int> iu_, iv_, wPlane_, cOffset_;
    long long cTaps_ = 0;
    int mpirank = 0, doSort = 0, runType = 0;

public:
    // the reference's members (current/tConvolveMPI/Benchmark.h:78-113)
    int nSamples = 0, wSize = 0, nChan = 1, gSize = 0;
    Coord uvCellSize = 0.0;
    Real baseline = 0.f;
    std::vector<Coord> u, v, w;
    std::vector<int> cOffset0, sSize;
    std::vector<long long> numPerPlane;
    int m_support = 0;
    int overSample = 0;
    Coord wCellSize = 0.0;
    uint64_t next;   // LCG state, seed 1
};

}  // namespace tcv

#endif
