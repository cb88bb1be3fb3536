/*
 * tcv_oracle.c -- CPU restatement of tConvolve's W-projection gridding path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * tcv_oracle.h for what pins each function.
 *
 * Arithmetic follows the reference statement by statement: the same types
 * (double Coord, float Real, std::complex<float> Value), the same evaluation
 * order, no FMA contraction (built with -ffp-contract=off, as the reference's
 * x86-64 build without -march produces no FMA), and the same libm calls.
 * Complex products use the limited-range formula the reference gets from
 * -fcx-limited-range (attic/tConvolveMPI/Makefile:24):
 *     (a+ib)(c+id) = (ac - bd) + i(ad + bc).
 *
 * Where the reference calls an unqualified sqrt/ceil on a float expression,
 * C++ overload resolution picks the double function (<cmath> only adds float
 * overloads inside namespace std), so the restatement promotes to double at
 * exactly those points.
 */
#define _GNU_SOURCE
#include "tcv_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MAXINT 2147483647u

/* a1 -- attic/tConvolveMPI/Benchmark.cc:53-58.  `next` is a 64-bit unsigned
 * long (LP64), so the product wraps at 2^64, not 2^32. */
int orc_random_int(uint64_t *next)
{
    *next = *next * 1103515245ull + 12345ull;
    return (int)((unsigned int)(*next / 65536ull) % MAXINT);
}

/* a2 -- attic/tConvolveMPI/Benchmark.cc:71-74.  `baseline / 2` is an integer
 * division in the reference (baseline is a const int). */
void orc_attic_uvw(int nSamples, int baseline, uint64_t *next,
                   double *u, double *v, double *w)
{
    for (int i = 0; i < nSamples; i++) {
        u[i] = baseline * (double)orc_random_int(next) / (double)MAXINT - baseline / 2;
        v[i] = baseline * (double)orc_random_int(next) / (double)MAXINT - baseline / 2;
        w[i] = baseline * (double)orc_random_int(next) / (double)MAXINT - baseline / 2;
    }
}

/* a2 -- attic/tConvolveMPI/Benchmark.cc:86-90 */
void orc_attic_freq(int nChan, double *freq)
{
    for (int i = 0; i < nChan; i++)
        freq[i] = (1.4e9 - 2.0e5 * (double)i / (double)nChan) / 2.998e8;
}

/* complex<float> *= complex<float>, limited range */
static inline void cmul_inplace(float *z, float br, float bi)
{
    float ar = z[0], ai = z[1];
    z[0] = ar * br - ai * bi;
    z[1] = ar * bi + ai * br;
}

/* a4 -- attic/tConvolveMPI/Benchmark.cc:206-270 */
void orc_attic_initC(double freq0, double cellSize, int wSize, int baseline,
                     int overSample, int *support, double *wCellSize, float *C)
{
    /* :212-213 */
    *support = (int)(1.5 * sqrt(abs(baseline) * cellSize * freq0) / cellSize);
    /* :217 */
    *wCellSize = 2 * baseline * freq0 / wSize;
    if (!C) return;

    const int sSize = 2 * (*support) + 1;
    const int cCenter = (sSize - 1) / 2;
    const long total = (long)sSize * sSize * overSample * overSample * wSize;

    /* :236-258 -- pow(x, 2) is exact-rounded x*x */
    for (int k = 0; k < wSize; k++) {
        double w = (double)(k - wSize / 2);
        double fScale = sqrt(fabs(w) * (*wCellSize) * freq0) / cellSize;
        for (int osj = 0; osj < overSample; osj++) {
            for (int osi = 0; osi < overSample; osi++) {
                for (int j = 0; j < sSize; j++) {
                    double tj = (double)(j - cCenter) + (double)osj / (double)overSample;
                    double j2 = tj * tj;
                    for (int i = 0; i < sSize; i++) {
                        double ti = (double)(i - cCenter) + (double)osi / (double)overSample;
                        double r2 = j2 + ti * ti;
                        long cind = i + (long)sSize * (j + (long)sSize * (osi + overSample * (osj + (long)overSample * k)));
                        C[2 * cind] = (w != 0.0) ? (float)cos(r2 / (w * fScale)) : (float)exp(-r2);
                        C[2 * cind + 1] = 0.0f;
                    }
                }
            }
        }
    }

    /* :260-269 -- float running sum of |C| (cabsf), then a float scale */
    float sumC = 0.0f;
    for (long i = 0; i < total; i++)
        sumC += cabsf(CMPLXF(C[2 * i], C[2 * i + 1]));
    const float scale = (float)(wSize * overSample * overSample) / sumC;
    for (long i = 0; i < total; i++)
        cmul_inplace(&C[2 * i], scale, 0.0f);
}

/* a6 -- attic/tConvolveMPI/Benchmark.cc:281-324 */
void orc_attic_initCOffset(int nSamples, const double *u, const double *v,
                           const double *w, int nChan, const double *freq,
                           double cellSize, double wCellSize, int wSize,
                           int gSize, int support, int overSample,
                           int *iu, int *iv, int *cOffset)
{
    const int sSize = 2 * support + 1;
    for (int i = 0; i < nSamples; i++) {
        for (int chan = 0; chan < nChan; chan++) {
            const int dind = i * nChan + chan;

            double uScaled = freq[chan] * u[i] / cellSize;
            int a = (int)uScaled;
            if (uScaled < (double)a) a -= 1;
            int fracu = (int)(overSample * (uScaled - (double)a));
            iu[dind] = a + gSize / 2;

            double vScaled = freq[chan] * v[i] / cellSize;
            int b = (int)vScaled;
            if (vScaled < (double)b) b -= 1;
            int fracv = (int)(overSample * (vScaled - (double)b));
            iv[dind] = b + gSize / 2;

            double wScaled = freq[chan] * w[i] / wCellSize;
            int woff = wSize / 2 + (int)wScaled;
            cOffset[dind] = sSize * sSize * (fracu + overSample * (fracv + overSample * woff));
        }
    }
}

/* a5 -- current/tConvolveMPI/Benchmark.cc:471-586 */
long orc_current_initC(double uvCellSize, int wSize, int support, int overSample,
                       double wCellSize, int *sSize, int *cOffset0, float *C)
{
    const int sSizeMax = 2 * support + 1;
    if (wSize == 1) sSize[0] = sSizeMax;                    /* :487-489 */

    long offsetCount = 0;
    for (int k = 0; k < wSize; k++) {
        const int wind = (int)(double)(k - wSize / 2);      /* :502 */
        const double w = wind * wCellSize;
        double fScale = 0.0;
        if (wind != 0) fScale = uvCellSize * uvCellSize / w;

        cOffset0[k] = (int)offsetCount;                      /* :509 */
        if (wSize > 1) {                                     /* :510-515 */
            const float wPart = (float)(w / uvCellSize / uvCellSize);
            const float aPart = 7.0f;
            sSize[k] = (int)ceil(sqrt((double)(aPart * aPart + wPart * wPart)));
            sSize[k] += (sSize[k] + 1) % 2;
        }
        const int s = sSize[k];
        const int cCenter = s / 2;
        const long planeTaps = (long)s * s * overSample * overSample;

        if (C) {
            double sumC = 0.0;
            for (int osj = 0; osj < overSample; osj++) {
                for (int osi = 0; osi < overSample; osi++) {
                    long osOffset = (long)s * s * (osi + overSample * osj) + offsetCount;
                    for (int j = 0; j < s; j++) {
                        double tj = (double)(j - cCenter) + (double)osj / (double)overSample;
                        double j2 = tj * tj;
                        for (int i = 0; i < s; i++) {
                            long cind = i + (long)s * j + osOffset;
                            double ti = (double)(i - cCenter) + (double)osi / (double)overSample;
                            double r2 = j2 + ti * ti;
                            float *z = &C[2 * cind];
                            z[0] = (float)exp(-r2);                 /* :535 */
                            z[1] = 0.0f;
                            if ((wind != 0) && (r2 < (double)(s / 2))) { /* :538-542 */
                                const float n = (float)sqrt(1. - r2 * fScale / w);
                                const float phase = (float)(-2. * 3.141593 * (r2 * fScale + w * ((double)n - 1.)));
                                /* Value(n/w) * Value(sin(phase), -cos(phase)) */
                                float pr = (float)((double)n / w), pi = 0.0f;
                                float qr = sinf(phase), qi = -cosf(phase);
                                float tr = pr * qr - pi * qi;
                                float ti2 = pr * qi + pi * qr;
                                cmul_inplace(z, tr, ti2);
                            }
                            sumC += (double)cabsf(CMPLXF(z[0], z[1]));  /* :544 */
                        }
                    }
                }
            }
            /* :552-555 */
            const float normC = (float)(overSample * overSample / sumC);
            for (long i = 0; i < planeTaps; i++)
                cmul_inplace(&C[2 * (i + offsetCount)], normC, 0.0f);
        }
        offsetCount += planeTaps;
    }
    return offsetCount;
}

/* a6 -- current/tConvolveMPI/Benchmark.cc:597-681 */
void orc_current_initCOffset(int nSamples, const double *u, const double *v,
                             const double *w, int nChan, const double *wavenumber,
                             double uvCellSize, double wCellSize, int wSize,
                             int gSize, int overSample, const int *sSize,
                             const int *cOffset0, int *iu, int *iv, int *wPlane,
                             int *cOffset, long *numPerPlane)
{
    for (int k = 0; k < wSize; k++) numPerPlane[k] = 0;
    for (int i = 0; i < nSamples; i++) {
        for (int chan = 0; chan < nChan; chan++) {
            const int dind = i * nChan + chan;

            double uScaled = wavenumber[chan] * u[i] / uvCellSize;
            int a = (int)uScaled;
            if (uScaled < (double)a) a -= 1;
            int fracu = (int)(overSample * (uScaled - (double)a));
            iu[dind] = a + gSize / 2;

            double vScaled = wavenumber[chan] * v[i] / uvCellSize;
            int b = (int)vScaled;
            if (vScaled < (double)b) b -= 1;
            int fracv = (int)(overSample * (vScaled - (double)b));
            iv[dind] = b + gSize / 2;

            int woff = 0;
            if (wCellSize > 0.0) {
                double wScaled = wavenumber[chan] * w[i] / wCellSize;
                woff = wSize / 2 + (int)wScaled;
            }
            wPlane[dind] = woff;
            cOffset[dind] = sSize[woff] * sSize[woff] * (fracu + overSample * fracv) + cOffset0[woff];
            numPerPlane[woff]++;
        }
    }
}

/* current/tConvolveMPI/Benchmark.cc:64-156 (Real is float; note the float
 * arithmetic and the double sqrt/ceil the unqualified calls resolve to). */
int orc_current_params(int runType, int *gSize, double *uvCellSize, int *wSize,
                       int *support, int *overSample, double *wCellSize,
                       float *baseline)
{
    const int apertureDiam = 12;
    const float maxFreqHz = 1.0e9f;
    const float lambda = (float)(2.998e8 / (double)maxFreqHz);
    float imgOS, imgExt, bl;
    if (runType == 0 || runType == 1) {
        float wmax, fov, wPart, aPart = 7.0f;
        int os;
        if (runType == 0) { bl = 6440.f; imgOS = 4.0f; imgExt = 3.0f; os = 4; }
        else              { bl = 2000.f; imgOS = 3.1f; imgExt = 2.0f; os = 8; }
        wmax = bl / lambda;
        fov = lambda / (float)apertureDiam * imgExt;
        wPart = wmax * fov * fov;
        *overSample = os;
        *support = (int)ceil(sqrt((double)(aPart * aPart + wPart * wPart))) / 2;
        int ws = (int)ceil((double)((float)os * wPart));
        ws += (ws + 1) % 2;
        *wSize = ws;
        *wCellSize = (double)(2.0f * wmax / (float)(ws - 1));
    } else if (runType == 2) {
        bl = 6440.f; imgOS = 2.0f; imgExt = 1.0f;
        *overSample = 1; *support = 0; *wSize = 1; *wCellSize = 0.0;
    } else if (runType == 3) {
        bl = 6440.f; imgOS = 4.0f; imgExt = 1.923f;
        *overSample = 128; *support = 3; *wSize = 1; *wCellSize = 0.0;
    } else if (runType == 4) {
        bl = 6440.f; imgOS = 4.0f; imgExt = 1.923f;
        *overSample = 8; *support = 43; *wSize = 1; *wCellSize = 0.0;
    } else {
        return 1;
    }
    *gSize = (int)ceil((double)(bl / (float)apertureDiam * imgOS * imgExt));   /* :154 */
    *uvCellSize = (double)(bl / lambda * imgOS / (float)(*gSize));           /* :156 */
    *baseline = bl;
    return 0;
}

/* ASKAP 36-antenna layout (metres east / north), the table of
 * current/tConvolveMPI/Benchmark.cc:180-197. */
static const double kEast[36] = {
    -42.43847222, -15.46047222, -6.48847222, -51.41747222, -116.43047222,
    93.22152778, 200.42152778, -80.24847222, -286.64847222, -138.75447222,
    225.51252778, 353.48652778, 396.28152778, -67.29847222, -782.13847222,
    -678.55347222, -539.25647222, -149.22347222, 175.37552778, 463.88152778,
    643.72952778, 803.40152778, -43.14647222, 36.03352778, -656.05547222,
    -435.77447222, -1112.94147222, 207.32652778, 523.49152778, 1186.61752778,
    2178.51052778, 2982.48652778, -17.41247222, -3017.44747222, -2213.43647222,
    -19.20647222};
static const double kNorth[36] = {
    -105.22933333, -118.24033333, -97.73933333, -70.22133333, -73.72633333,
    6.77066667, -215.20933333, -343.73933333, 5.32066667, 174.26666667,
    235.30966667, 164.24666667, -469.23433333, -565.23833333, -263.22233333,
    260.21066667, 417.28966667, 270.27966667, 376.29066667, 209.79766667,
    216.77666667, 230.75266667, -762.23633333, -1083.75233333, 548.27766667,
    562.27366667, 835.74566667, 1093.28166667, 647.98066667, 693.23866667,
    887.75566667, -2612.27533333, -2916.19433333, -2112.20733333, 887.76166667,
    3084.83866667};

/* current/tConvolveMPI/Benchmark.cc:168-258 */
int orc_current_uvw(float baseline, uint64_t *next, double *u, double *v, double *w)
{
    const double obslen = 12., scanlen = 5.;
    const int nScans = (int)(obslen * 3600. / scanlen);
    const double lat = -26.6970 * 3.141593 / 180.0;
    const double dec = lat;
    const double cdec = cos(dec), sdec = sin(dec);
    double X[36], Y[36], Z[36], BX[630], BY[630], BZ[630];
    for (int i = 0; i < 36; i++) {
        X[i] = -kNorth[i] * sin(lat);
        Y[i] = kEast[i];
        Z[i] = kNorth[i] * cos(lat);
    }
    int bl = 0;
    for (int i = 0; i < 35; i++) {
        for (int j = i + 1; j < 36; j++) {
            double dX = X[i] - X[j], dY = Y[i] - Y[j], dZ = Z[i] - Z[j];
            if (dX * dX + dY * dY + dZ * dZ > (double)(baseline * baseline)) continue;
            BX[bl] = dX; BY[bl] = dY; BZ[bl] = dZ;
            bl++;
        }
    }
    const int nBaselines = bl;
    const int nSamples = nScans * nBaselines;
    if (!u) return nSamples;
    for (int i = 0; i < nSamples; i++) {
        const int b = (int)(nBaselines * ((double)orc_random_int(next) / (double)MAXINT));
        const double ha = obslen * 3.141593 / 12.0 * (((double)orc_random_int(next) / (double)MAXINT) - 0.5);
        const double cha = cos(ha), sha = sin(ha);
        u[i] = sha * BX[b] + cha * BY[b];
        v[i] = -sdec * cha * BX[b] + sdec * sha * BY[b] + cdec * BZ[b];
        w[i] = cdec * cha * BX[b] - cdec * sha * BY[b] + sdec * BZ[b];
    }
    return nSamples;
}

static inline int vis_size(int d, const int *wPlane, const int *sSize, int sUniform)
{
    return wPlane ? sSize[wPlane[d]] : sUniform;
}

/* a7 -- attic/tConvolveMPI/Benchmark.cc:125-155; current :373-413.
 * grid += d * C, one complex product then one complex add per tap. */
void orc_grid(int nVis, const float *data, const int *iu, const int *iv,
              const int *cOffset, const int *wPlane, const int *sSize,
              int sSizeUniform, const float *C, float *grid, int gSize)
{
    for (int dind = 0; dind < nVis; ++dind) {
        const int s = vis_size(dind, wPlane, sSize, sSizeUniform);
        const int support = s / 2;
        long gind = iu[dind] + (long)gSize * iv[dind] - support;
        long cind = cOffset[dind];
        const float dr = data[2 * dind], di = data[2 * dind + 1];
        for (int suppv = 0; suppv < s; suppv++) {
            float *g = &grid[2 * gind];
            const float *c = &C[2 * cind];
            for (int suppu = 0; suppu < s; suppu++) {
                const float cr = c[2 * suppu], ci = c[2 * suppu + 1];
                const float pr = dr * cr - di * ci;
                const float pi = dr * ci + di * cr;
                g[2 * suppu] += pr;
                g[2 * suppu + 1] += pi;
            }
            gind += gSize;
            cind += s;
        }
    }
}

/* a8 -- attic/tConvolveMPI/Benchmark.cc:158-195; current :416-460.
 * data = sum over the footprint of grid * C, sequential float accumulation. */
void orc_degrid(int nVis, const float *grid, int gSize, const int *iu,
                const int *iv, const int *cOffset, const int *wPlane,
                const int *sSize, int sSizeUniform, const float *C, float *data)
{
    for (int dind = 0; dind < nVis; ++dind) {
        const int s = vis_size(dind, wPlane, sSize, sSizeUniform);
        const int support = s / 2;
        long gind = iu[dind] + (long)gSize * iv[dind] - support;
        long cind = cOffset[dind];
        float re = 0.0f, im = 0.0f;
        for (int suppv = 0; suppv < s; suppv++) {
            const float *g = &grid[2 * gind];
            const float *c = &C[2 * cind];
            for (int suppu = 0; suppu < s; suppu++) {
                const float gr = g[2 * suppu], gi = g[2 * suppu + 1];
                const float cr = c[2 * suppu], ci = c[2 * suppu + 1];
                re += gr * cr - gi * ci;
                im += gr * ci + gi * cr;
            }
            gind += gSize;
            cind += s;
        }
        data[2 * dind] = re;
        data[2 * dind + 1] = im;
    }
}

void orc_grid_f64(int nVis, const double *data, const int *iu, const int *iv,
                  const int *cOffset, const int *wPlane, const int *sSize,
                  int sSizeUniform, const float *C, double *grid, int gSize)
{
    for (int dind = 0; dind < nVis; ++dind) {
        const int s = vis_size(dind, wPlane, sSize, sSizeUniform);
        long gind = iu[dind] + (long)gSize * iv[dind] - s / 2;
        long cind = cOffset[dind];
        const double dr = data[2 * dind], di = data[2 * dind + 1];
        for (int suppv = 0; suppv < s; suppv++) {
            double *g = &grid[2 * gind];
            const float *c = &C[2 * cind];
            for (int suppu = 0; suppu < s; suppu++) {
                const double cr = c[2 * suppu], ci = c[2 * suppu + 1];
                g[2 * suppu] += dr * cr - di * ci;
                g[2 * suppu + 1] += dr * ci + di * cr;
            }
            gind += gSize;
            cind += s;
        }
    }
}

void orc_degrid_f64(int nVis, const double *grid, int gSize, const int *iu,
                    const int *iv, const int *cOffset, const int *wPlane,
                    const int *sSize, int sSizeUniform, const float *C,
                    double *data)
{
    for (int dind = 0; dind < nVis; ++dind) {
        const int s = vis_size(dind, wPlane, sSize, sSizeUniform);
        long gind = iu[dind] + (long)gSize * iv[dind] - s / 2;
        long cind = cOffset[dind];
        double re = 0.0, im = 0.0;
        for (int suppv = 0; suppv < s; suppv++) {
            const double *g = &grid[2 * gind];
            const float *c = &C[2 * cind];
            for (int suppu = 0; suppu < s; suppu++) {
                const double gr = g[2 * suppu], gi = g[2 * suppu + 1];
                const double cr = c[2 * suppu], ci = c[2 * suppu + 1];
                re += gr * cr - gi * ci;
                im += gr * ci + gi * cr;
            }
            gind += gSize;
            cind += s;
        }
        data[2 * dind] = re;
        data[2 * dind + 1] = im;
    }
}

/* a7 restricted to a window of the grid: the cells [x0, x0 + w) x [y0, y0 + h)
 * of gridKernel's result (attic/tConvolveMPI/Benchmark.cc:125-155), into a
 * window-sized h x w complex array (row y - y0, column x - x0) that it
 * accumulates into.  Visibilities are taken in index order and every cell
 * sums them in that order with the same float operations, so each window
 * value equals orc_grid's value of that cell bit for bit; only the footprint
 * rows and columns inside the window are visited, which makes a window of a
 * full-size job cheap.  f64 != 0: double accumulation into a double window
 * (orc_grid_f64's values). */
void orc_grid_window(int nVis, const float *data, const int *iu, const int *iv,
                     const int *cOffset, const int *wPlane, const int *sSize,
                     int sSizeUniform, const float *C, int x0, int y0, int w, int h,
                     void *win, int f64)
{
    for (int dind = 0; dind < nVis; ++dind) {
        const int s = vis_size(dind, wPlane, sSize, sSizeUniform);
        const int u0 = iu[dind] - s / 2, v0 = iv[dind];
        /* footprint [u0, u0 + s) x [v0, v0 + s) against the window */
        const int ua = u0 > x0 ? u0 : x0, ub = (u0 + s < x0 + w) ? u0 + s : x0 + w;
        const int va = v0 > y0 ? v0 : y0, vb = (v0 + s < y0 + h) ? v0 + s : y0 + h;
        if (ua >= ub || va >= vb) continue;
        const float dr = data[2 * dind], di = data[2 * dind + 1];
        for (int y = va; y < vb; y++) {
            const float *c = &C[2 * ((long)cOffset[dind] + (long)(y - v0) * s + (ua - u0))];
            const long o = 2 * ((long)(y - y0) * w + (ua - x0));
            if (f64) {
                double *g = (double *)win + o;
                const double ddr = dr, ddi = di;
                for (int x = 0; x < ub - ua; x++) {
                    const double cr = c[2 * x], ci = c[2 * x + 1];
                    g[2 * x] += ddr * cr - ddi * ci;
                    g[2 * x + 1] += ddr * ci + ddi * cr;
                }
            } else {
                float *g = (float *)win + o;
                for (int x = 0; x < ub - ua; x++) {
                    const float cr = c[2 * x], ci = c[2 * x + 1];
                    const float pr = dr * cr - di * ci;
                    const float pi = dr * ci + di * cr;
                    g[2 * x] += pr;
                    g[2 * x + 1] += pi;
                }
            }
        }
    }
}

/* current/tConvolveMPI/Benchmark.cc:275-298: counting sort of the index
 * arrays by w-plane (setSort(1)): plane k's visibilities go, in their input
 * order, to the positions starting at the sum of numPerPlane over the planes
 * before k.  Like the reference, the visibility data is not permuted. */
int orc_sort_by_plane(int n, int wSize, const long long *numPerPlane, int *iu, int *iv,
                      int *wPlane, int *cOffset)
{
    int *tmp = (int *)malloc(sizeof(int) * 4 * (size_t)(n > 0 ? n : 1));
    long long *next = (long long *)malloc(sizeof(long long) * (size_t)(wSize > 0 ? wSize : 1));
    if (!tmp || !next) {
        free(tmp);
        free(next);
        return -1;
    }
    memcpy(tmp, iu, sizeof(int) * (size_t)n);
    memcpy(tmp + n, iv, sizeof(int) * (size_t)n);
    memcpy(tmp + 2 * (size_t)n, wPlane, sizeof(int) * (size_t)n);
    memcpy(tmp + 3 * (size_t)n, cOffset, sizeof(int) * (size_t)n);
    next[0] = 0;
    for (int k = 1; k < wSize; k++) next[k] = next[k - 1] + numPerPlane[k - 1];
    for (int i = 0; i < n; i++) {
        const int k = tmp[2 * (size_t)n + i];
        const long long j = next[k]++;
        iu[j] = tmp[i];
        iv[j] = tmp[(size_t)n + i];
        wPlane[j] = k;
        cOffset[j] = tmp[3 * (size_t)n + i];
    }
    free(tmp);
    free(next);
    return 0;
}

long long orc_grid_points(int nVis, const int *wPlane, const int *sSize,
                          int sSizeUniform)
{
    long long n = 0;
    for (int d = 0; d < nVis; d++) {
        const long long s = vis_size(d, wPlane, sSize, sSizeUniform);
        n += s * s;
    }
    return n;
}

/* ---- a11: CPU baseline runner ------------------------------------------ */

typedef struct {
    int t, nVis, nInput, gSize, sUniform;
    const float *data, *C;
    const int *iu, *iv, *cOffset, *wPlane, *sSize;
    float *grid, *out, *sdata;
    int *siu, *siv, *scoff, *swp;
    pthread_barrier_t *bar;
    int ok;
} worker_t;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void *worker_main(void *arg)
{
    worker_t *wk = (worker_t *)arg;
    const long g2 = (long)wk->gSize * wk->gSize;
    wk->grid = (float *)calloc((size_t)g2 * 2, sizeof(float));
    wk->out = (float *)calloc((size_t)wk->nVis * 2, sizeof(float));
    wk->sdata = (float *)malloc((size_t)wk->nVis * 2 * sizeof(float));
    wk->siu = (int *)malloc((size_t)wk->nVis * sizeof(int));
    wk->siv = (int *)malloc((size_t)wk->nVis * sizeof(int));
    wk->scoff = (int *)malloc((size_t)wk->nVis * sizeof(int));
    wk->swp = wk->wPlane ? (int *)malloc((size_t)wk->nVis * sizeof(int)) : NULL;
    wk->ok = wk->grid && wk->out && wk->sdata && wk->siu && wk->siv && wk->scoff &&
             (!wk->wPlane || wk->swp);
    if (wk->ok) {
        for (int k = 0; k < wk->nVis; k++) {
            const long src = ((long)wk->t * wk->nVis + k) % wk->nInput;
            wk->sdata[2 * k] = wk->data[2 * src];
            wk->sdata[2 * k + 1] = wk->data[2 * src + 1];
            wk->siu[k] = wk->iu[src];
            wk->siv[k] = wk->iv[src];
            wk->scoff[k] = wk->cOffset[src];
            if (wk->swp) wk->swp[k] = wk->wPlane[src];
        }
    }
    pthread_barrier_wait(wk->bar);   /* set-up done */
    pthread_barrier_wait(wk->bar);   /* start gridding */
    if (wk->ok)
        orc_grid(wk->nVis, wk->sdata, wk->siu, wk->siv, wk->scoff, wk->swp,
                 wk->sSize, wk->sUniform, wk->C, wk->grid, wk->gSize);
    pthread_barrier_wait(wk->bar);   /* gridding done */
    pthread_barrier_wait(wk->bar);   /* start degridding */
    if (wk->ok)
        orc_degrid(wk->nVis, wk->grid, wk->gSize, wk->siu, wk->siv, wk->scoff,
                   wk->swp, wk->sSize, wk->sUniform, wk->C, wk->out);
    pthread_barrier_wait(wk->bar);   /* degridding done */
    free(wk->grid); free(wk->out); free(wk->sdata);
    free(wk->siu); free(wk->siv); free(wk->scoff); free(wk->swp);
    return NULL;
}

int orc_cpu_baseline(int nThreads, int nVis, int nInput, const float *data,
                     const int *iu, const int *iv, const int *cOffset,
                     const int *wPlane, const int *sSize, int sSizeUniform,
                     const float *C, int gSize, double *tGrid, double *tDegrid)
{
    if (nThreads < 1 || nVis < 1 || nInput < 1) return 1;
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)nThreads + 1);
    worker_t *wk = (worker_t *)calloc((size_t)nThreads, sizeof(worker_t));
    pthread_t *th = (pthread_t *)calloc((size_t)nThreads, sizeof(pthread_t));
    for (int t = 0; t < nThreads; t++) {
        worker_t w0 = {t, nVis, nInput, gSize, sSizeUniform, data, C, iu, iv,
                       cOffset, wPlane, sSize, NULL, NULL, NULL, NULL, NULL,
                       NULL, NULL, &bar, 0};
        wk[t] = w0;
        pthread_create(&th[t], NULL, worker_main, &wk[t]);
    }
    pthread_barrier_wait(&bar);
    pthread_barrier_wait(&bar);
    double t0 = now_s();
    pthread_barrier_wait(&bar);
    double t1 = now_s();
    pthread_barrier_wait(&bar);
    double t2 = now_s();
    pthread_barrier_wait(&bar);
    double t3 = now_s();
    int ok = 1;
    for (int t = 0; t < nThreads; t++) {
        pthread_join(th[t], NULL);
        ok &= wk[t].ok;
    }
    pthread_barrier_destroy(&bar);
    free(wk);
    free(th);
    *tGrid = t1 - t0;
    *tDegrid = t3 - t2;
    return ok ? 0 : 2;
}

/* ---- major cycle --------------------------------------------------------- */

/* current/tMajorACC/tMajorACC.cc:620-627 */
void orc_fft_fix(float *grid, long long nPixels, float scale)
{
    for (long long i = 0; i < nPixels; i++) {
        grid[2 * i] = grid[2 * i] * scale;
        grid[2 * i + 1] = 0.0f;
    }
}

/* current/tMajorACC/tMajorACC.cc:689-701 */
void orc_find_peak(const float *image, long long nPixels, float *maxVal,
                   long long *maxPos)
{
    float mv = 0.0f;
    long long mp = 0;
    for (long long i = 0; i < nPixels; i++) {
        const float v = image[2 * i];
        if (fabsf(v) > fabsf(mv)) {
            mv = v;
            mp = i;
        }
    }
    *maxVal = mv;
    *maxPos = mp;
}

/* current/tMajorACC/tMajorACC.cc:723-750 (idxToPos / posToIdx :677-687) */
void orc_subtract_psf(const float *psf, long long psfWidth, float *residual,
                      long long residualWidth, long long peakPos,
                      long long psfPeakPos, float absPeakVal, float gain)
{
    const long long rx = peakPos % residualWidth, ry = peakPos / residualWidth;
    const long long px = psfPeakPos % psfWidth, py = psfPeakPos / psfWidth;
    const long long diffx = rx - px;
    const long long diffy = ry - px;   /* as in the reference (not ry - py) */
    const long long startx = rx - px > 0 ? rx - px : 0;
    const long long starty = ry - py > 0 ? ry - py : 0;
    long long stopx = rx + (psfWidth - px - 1), stopy = ry + (psfWidth - py - 1);
    if (stopx > residualWidth - 1) stopx = residualWidth - 1;
    if (stopy > residualWidth - 1) stopy = residualWidth - 1;
    const float amp = gain * absPeakVal;
    for (long long y = starty; y <= stopy; ++y) {
        const long long sy = y - diffy;
        if (sy < 0 || sy >= psfWidth) continue;
        for (long long x = startx; x <= stopx; ++x) {
            float *r = residual + 2 * (y * residualWidth + x);
            r[0] = r[0] - amp * psf[2 * (sy * psfWidth + (x - diffx))];
        }
    }
}

/* current/tMajorACC/tMajorACC.cc:800-840 */
void orc_deconvolve(float *residual, long long dirtyWidth, const float *psf,
                    long long psfWidth, float *model, int niters, float gain,
                    float *psfPeakVal, long long *psfPeakPos,
                    float *dirtyPeakVal, long long *dirtyPeakPos)
{
    float pv;
    long long pp;
    orc_find_peak(psf, psfWidth * psfWidth, &pv, &pp);
    if (psfPeakVal) *psfPeakVal = pv;
    if (psfPeakPos) *psfPeakPos = pp;
    for (int i = 0; i < niters; ++i) {
        float av;
        long long ap;
        orc_find_peak(residual, dirtyWidth * dirtyWidth, &av, &ap);
        if (i == 0) {
            if (dirtyPeakVal) *dirtyPeakVal = av;
            if (dirtyPeakPos) *dirtyPeakPos = ap;
        }
        model[2 * ap] = model[2 * ap] + av * gain;
        orc_subtract_psf(psf, psfWidth, residual, dirtyWidth, ap, pp, av, gain);
    }
}

/* current/tMajorACC/tMajorACC.cc:1549-1551 */
void orc_vis_subtract(float *data, const float *model, long long nVis)
{
    for (long long i = 0; i < 2 * nVis; i++) data[i] = data[i] - model[i];
}
