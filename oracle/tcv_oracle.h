/*
 * tcv_oracle.h -- CPU restatement of tConvolve's W-projection gridding path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it, and never as the
 * thing that is measured or shipped.  The product path (libtcv.so) never links
 * or calls it.
 *
 * Complex values are interleaved float pairs (re, im), exactly the memory
 * layout of std::complex<float> used by the reference.  Every function cites
 * the reference file:line it restates (paths relative to the reference root).
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_random_int        closed-form first draws + compiled-reference fixture
 *   orc_attic_uvw         compiled-reference fixture (first samples, checksums)
 *   orc_attic_initC       printed scalars (support 64, wCellSize 566.034,
 *                         shape [129,129,8,8,33]) + closed form per tap +
 *                         compiled-reference checksum
 *   orc_current_initC     compiled-reference checksum (runType 0..4), per-plane
 *                         normalisation invariant sum|C| = os^2
 *   orc_*_initCOffset     compiled-reference fixture, FP64 closed form,
 *                         range invariants
 *   orc_grid / orc_degrid single-visibility patch == C slice, linearity,
 *                         bilinear adjointness (FP64), grid==1 => sum of slice,
 *                         brute force in FP64, compiled-reference checksums
 *   orc_sort_by_plane     stable (input order kept within a plane), planes
 *                         contiguous in order, a permutation of the input
 *   orc_grid_window       bit-identical to orc_grid / orc_grid_f64 on every
 *                         window cell (windows on edges, corners, straddling),
 *                         hence to the compiled reference's grid
 *   orc_find_peak         numpy argmax of |re| (first occurrence), ties, zeros
 *   orc_subtract_psf      delta PSF changes one pixel by gain*peak; all-ones
 *                         PSF changes exactly the overlap rectangle; the
 *                         reference's row offset with an off-diagonal peak
 *   orc_deconvolve        closed form on a delta PSF (peak * 0.9^n), model sum
 *                         + residual conservation on a delta PSF
 *   orc_fft_fix, orc_vis_subtract   element-wise closed forms
 * (tMajorACC needs FFTW and an OpenACC compiler, neither present here, so no
 * compiled reference backs the major-cycle functions.)
 */
#ifndef TCV_ORACLE_H
#define TCV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* a1: Benchmark::randomInt -- attic/tConvolveMPI/Benchmark.cc:53-58 */
int orc_random_int(uint64_t *next);

/* a2: Benchmark::init (u, v, w part) -- attic/tConvolveMPI/Benchmark.cc:69-80 */
void orc_attic_uvw(int nSamples, int baseline, uint64_t *next,
                   double *u, double *v, double *w);

/* a2: channel frequencies -- attic/tConvolveMPI/Benchmark.cc:86-90 */
void orc_attic_freq(int nChan, double *freq);

/* a4: Benchmark::initC (attic, real C, uniform size) --
 * attic/tConvolveMPI/Benchmark.cc:206-270.  C must hold
 * sSize*sSize*overSample*overSample*wSize complex values; pass C == NULL to
 * only obtain support and wCellSize. */
void orc_attic_initC(double freq0, double cellSize, int wSize, int baseline,
                     int overSample, int *support, double *wCellSize, float *C);

/* a6: Benchmark::initCOffset (attic) -- attic/tConvolveMPI/Benchmark.cc:281-324 */
void orc_attic_initCOffset(int nSamples, const double *u, const double *v,
                           const double *w, int nChan, const double *freq,
                           double cellSize, double wCellSize, int wSize,
                           int gSize, int support, int overSample,
                           int *iu, int *iv, int *cOffset);

/* a5: Benchmark::initC (current, complex C, per-plane size) --
 * current/tConvolveMPI/Benchmark.cc:471-586.  sSize[] (wSize entries) is an
 * input when wSize == 1 (set to 2*support+1) and an output otherwise.
 * Returns the number of complex taps; C may be NULL to size the table. */
long orc_current_initC(double uvCellSize, int wSize, int support, int overSample,
                       double wCellSize, int *sSize, int *cOffset0, float *C);

/* a6: Benchmark::initCOffset (current) -- current/tConvolveMPI/Benchmark.cc:597-681 */
void orc_current_initCOffset(int nSamples, const double *u, const double *v,
                             const double *w, int nChan, const double *wavenumber,
                             double uvCellSize, double wCellSize, int wSize,
                             int gSize, int overSample, const int *sSize,
                             const int *cOffset0, int *iu, int *iv, int *wPlane,
                             int *cOffset, long *numPerPlane);

/* a5/a2: current-generation problem parameters for runType 0..4 --
 * current/tConvolveMPI/Benchmark.cc:64-156.  Returns 0 on success. */
int orc_current_params(int runType, int *gSize, double *uvCellSize, int *wSize,
                       int *support, int *overSample, double *wCellSize,
                       float *baseline);

/* a2: current-generation ASKAP-layout uvw generator --
 * current/tConvolveMPI/Benchmark.cc:168-258.  Returns nSamples; u/v/w may be
 * NULL to size the arrays (the LCG state is then left untouched). */
int orc_current_uvw(float baseline, uint64_t *next, double *u, double *v,
                    double *w);

/* a7: gridKernel -- attic/tConvolveMPI/Benchmark.cc:125-155 and
 * current/tConvolveMPI/Benchmark.cc:373-413.  Per visibility the kernel size is
 * sSize[wPlane[d]] when wPlane != NULL, else sSizeUniform.  Accumulates. */
void orc_grid(int nVis, const float *data, const int *iu, const int *iv,
              const int *cOffset, const int *wPlane, const int *sSize,
              int sSizeUniform, const float *C, float *grid, int gSize);

/* a8: degridKernel -- attic/tConvolveMPI/Benchmark.cc:158-195 and
 * current/tConvolveMPI/Benchmark.cc:416-460.  Overwrites data. */
void orc_degrid(int nVis, const float *grid, int gSize, const int *iu,
                const int *iv, const int *cOffset, const int *wPlane,
                const int *sSize, int sSizeUniform, const float *C, float *data);

/* FP64 oracle: same indices, float C promoted to double, double accumulation. */
void orc_grid_f64(int nVis, const double *data, const int *iu, const int *iv,
                  const int *cOffset, const int *wPlane, const int *sSize,
                  int sSizeUniform, const float *C, double *grid, int gSize);
void orc_degrid_f64(int nVis, const double *grid, int gSize, const int *iu,
                    const int *iv, const int *cOffset, const int *wPlane,
                    const int *sSize, int sSizeUniform, const float *C,
                    double *data);

/* a7 on a window [x0, x0 + w) x [y0, y0 + h) of the grid only, into an h x w
 * complex window (float, or double when f64 != 0) that it accumulates into:
 * the same values as orc_grid (orc_grid_f64) on those cells, bit for bit. */
void orc_grid_window(int nVis, const float *data, const int *iu, const int *iv,
                     const int *cOffset, const int *wPlane, const int *sSize,
                     int sSizeUniform, const float *C, int x0, int y0, int w, int h,
                     void *win, int f64);

/* a13: setSort(1) -- current/tConvolveMPI/Benchmark.cc:275-298: counting sort of
 * iu, iv, wPlane, cOffset by w-plane (stable), in place.  0 on success. */
int orc_sort_by_plane(int n, int wSize, const long long *numPerPlane, int *iu, int *iv,
                      int *wPlane, int *cOffset);

/* a10: grid points = sum over visibilities of sSize(w)^2 --
 * current/tConvolveMPI/Benchmark.cc:683-700 */
long long orc_grid_points(int nVis, const int *wPlane, const int *sSize,
                          int sSizeUniform);

/* a11: CPU baseline with the tConvolveMPI / tConvolveMT methodology --
 * attic/tConvolveMPI/tConvolveMPI.cc:82-122, attic/tConvolveMT/tConvolveMT.cc:89-117.
 * nThreads workers each grid and then degrid their own nVis visibilities into
 * their own gSize^2 grid (C is shared read-only), with a barrier before and
 * after each timed phase.  Visibility k of worker t is input index
 * (t*nVis + k) % nInput.  Times are steady-clock seconds.  Returns 0 on success. */
int orc_cpu_baseline(int nThreads, int nVis, int nInput, const float *data,
                     const int *iu, const int *iv, const int *cOffset,
                     const int *wPlane, const int *sSize, int sSizeUniform,
                     const float *C, int gSize, double *tGrid, double *tDegrid);

/* ---- major cycle (current/tMajorACC/tMajorACC.cc) ---------------------- */

/* fftFix: grid[i] = real(grid[i]) * scale (imaginary part set to 0) --
 * current/tMajorACC/tMajorACC.cc:620-627 */
void orc_fft_fix(float *grid, long long nPixels, float scale);

/* findPeak: the first pixel of largest |real part|; *maxVal is that real part
 * (signed), 0 and position 0 for an all-zero image --
 * current/tMajorACC/tMajorACC.cc:689-701 */
void orc_find_peak(const float *image, long long nPixels, float *maxVal,
                   long long *maxPos);

/* subtractPsf: residual -= gain * peak * psf (real parts), psf placed with
 * its peak on the residual peak; the row offset uses the reference's
 * diffy = ry - px (tMajorACC.cc:735); PSF rows that this offset puts outside
 * the PSF are skipped (the reference would read outside its buffer) --
 * current/tMajorACC/tMajorACC.cc:723-750 */
void orc_subtract_psf(const float *psf, long long psfWidth, float *residual,
                      long long residualWidth, long long peakPos,
                      long long psfPeakPos, float absPeakVal, float gain);

/* deconvolve: Hogbom CLEAN, niters iterations of findPeak / model update /
 * subtractPsf with the given gain (0.1 in the reference) --
 * current/tMajorACC/tMajorACC.cc:800-840.  Optional outputs: the PSF peak and
 * the first residual peak (the values the reference prints). */
void orc_deconvolve(float *residual, long long dirtyWidth, const float *psf,
                    long long psfWidth, float *model, int niters, float gain,
                    float *psfPeakVal, long long *psfPeakPos,
                    float *dirtyPeakVal, long long *dirtyPeakPos);

/* visibility update: data[i] = data[i] - model[i] --
 * current/tMajorACC/tMajorACC.cc:1549-1551 */
void orc_vis_subtract(float *data, const float *model, long long nVis);

#ifdef __cplusplus
}
#endif
#endif
