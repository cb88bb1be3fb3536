// Host set-up of the gridding problem: the reference's LCG, uvw generators,
// problem parameters and C(w) tables.  Compiled with g++ and the reference's
// flags (-O3 -fstrict-aliasing -fcx-limited-range, plus -ffp-contract=off) so
// every value, including compile-time folded libm calls, rounds exactly as in
// the reference build.  No GPU code here.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <thread>
#include <vector>

#include "tcv.h"

namespace {

static unsigned host_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    return std::max(1u, std::min(n, 32u));
}

template <class F>
static void parallel_for(int n, F f)
{
    const int nt = (int)std::min<unsigned>(host_threads(), (unsigned)std::max(n, 1));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([=]() {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

// ---- host set-up -----------------------------------------------------------

int tcv_random_int(uint64_t *next)
{
    // attic/tConvolveMPI/Benchmark.cc:53-58: 64-bit state, K&R constants.
    const unsigned int maxint = 2147483647u;
    *next = *next * 1103515245ull + 12345ull;
    return (int)((unsigned int)(*next / 65536ull) % maxint);
}

void tcv_attic_uvw(int nSamples, int baseline, uint64_t *next, double *u, double *v, double *w)
{
    // attic/tConvolveMPI/Benchmark.cc:71-74 (baseline / 2 is integer division)
    const double maxint = 2147483647.0;
    for (int i = 0; i < nSamples; i++) {
        u[i] = baseline * double(tcv_random_int(next)) / maxint - baseline / 2;
        v[i] = baseline * double(tcv_random_int(next)) / maxint - baseline / 2;
        w[i] = baseline * double(tcv_random_int(next)) / maxint - baseline / 2;
    }
}

tcv_status tcv_attic_initC(double freq0, double cellSize, int wSize, int baseline, int overSample,
                           int *support, double *wCellSize, float *C)
{
    // attic/tConvolveMPI/Benchmark.cc:206-270
    if (!support || !wCellSize || wSize < 1 || overSample < 1 || cellSize <= 0)
        return TCV_ERR_INVALID;
    *support = static_cast<int>(1.5 * std::sqrt(std::abs(baseline) * cellSize * freq0) / cellSize);
    *wCellSize = 2 * baseline * freq0 / wSize;
    if (!C) return TCV_OK;
    const int sSize = 2 * *support + 1;
    const int cCenter = (sSize - 1) / 2;
    const double wcs = *wCellSize;
    const long plane = (long)sSize * sSize * overSample * overSample;
    std::complex<float> *Cv = reinterpret_cast<std::complex<float> *>(C);
    // taps: one task per (plane, osj) slab, independent
    parallel_for(wSize * overSample, [&](int task) {
        const int k = task / overSample, osj = task % overSample;
        const double w = double(k - wSize / 2);
        const double fScale = std::sqrt(std::abs(w) * wcs * freq0) / cellSize;
        for (int osi = 0; osi < overSample; osi++)
            for (int j = 0; j < sSize; j++) {
                const double tj = double(j - cCenter) + double(osj) / double(overSample);
                const double j2 = tj * tj;
                for (int i = 0; i < sSize; i++) {
                    const double ti = double(i - cCenter) + double(osi) / double(overSample);
                    const double r2 = j2 + ti * ti;
                    const long cind = i + (long)sSize * (j + (long)sSize * (osi + overSample * (osj + (long)overSample * k)));
                    Cv[cind] = std::complex<float>(
                        w != 0.0 ? (float)std::cos(r2 / (w * fScale)) : (float)std::exp(-r2), 0.0f);
                }
            }
    });
    // normalisation: sequential float sum of |C| (the reference's order), then scale
    const long total = plane * wSize;
    float sumC = 0.0f;
    for (long i = 0; i < total; i++) sumC += hypotf(Cv[i].real(), Cv[i].imag());
    const float scale = float(wSize * overSample * overSample) / sumC;
    parallel_for(wSize, [&](int k) {
        for (long i = k * plane; i < (k + 1) * plane; i++) {
            const float re = Cv[i].real(), im = Cv[i].imag();
            Cv[i] = std::complex<float>(re * scale - im * 0.0f, re * 0.0f + im * scale);
        }
    });
    return TCV_OK;
}

tcv_status tcv_current_params(int runType, int *gSize, double *uvCellSize, int *wSize,
                              int *support, int *overSample, double *wCellSize, float *baseline)
{
    // current/tConvolveMPI/Benchmark.cc:64-156 (float Real arithmetic; the
    // unqualified sqrt/ceil there are the double functions)
    if (!gSize || !uvCellSize || !wSize || !support || !overSample || !wCellSize || !baseline)
        return TCV_ERR_INVALID;
    const int apertureDiam = 12;
    const float maxFreqHz = 1.0e9f;
    const float lambda = (float)(2.998e8 / (double)maxFreqHz);
    float imgOS, imgExt, bl;
    switch (runType) {
    case 0:
    case 1: {
        int os;
        if (runType == 0) { bl = 6440.f; imgOS = 4.0f; imgExt = 3.0f; os = 4; }
        else { bl = 2000.f; imgOS = 3.1f; imgExt = 2.0f; os = 8; }
        const float wmax = bl / lambda;
        const float fov = lambda / (float)apertureDiam * imgExt;
        const float wPart = wmax * fov * fov;
        const float aPart = 7.0f;
        *overSample = os;
        *support = (int)std::ceil(std::sqrt((double)(aPart * aPart + wPart * wPart))) / 2;
        int ws = (int)std::ceil((double)((float)os * wPart));
        ws += (ws + 1) % 2;
        *wSize = ws;
        *wCellSize = (double)(2.0f * wmax / (float)(ws - 1));
        break;
    }
    case 2: bl = 6440.f; imgOS = 2.0f; imgExt = 1.0f; *overSample = 1; *support = 0; *wSize = 1; *wCellSize = 0.0; break;
    case 3: bl = 6440.f; imgOS = 4.0f; imgExt = 1.923f; *overSample = 128; *support = 3; *wSize = 1; *wCellSize = 0.0; break;
    case 4: bl = 6440.f; imgOS = 4.0f; imgExt = 1.923f; *overSample = 8; *support = 43; *wSize = 1; *wCellSize = 0.0; break;
    default: return TCV_ERR_INVALID;
    }
    *gSize = (int)std::ceil((double)(bl / (float)apertureDiam * imgOS * imgExt));
    *uvCellSize = (double)(bl / lambda * imgOS / (float)(*gSize));
    *baseline = bl;
    return TCV_OK;
}

// ASKAP antenna positions (east, north metres): the layout table of
// current/tConvolveMPI/Benchmark.cc:180-197.
static const double kAskapEast[36] = {
    -42.43847222, -15.46047222, -6.48847222, -51.41747222, -116.43047222, 93.22152778,
    200.42152778, -80.24847222, -286.64847222, -138.75447222, 225.51252778, 353.48652778,
    396.28152778, -67.29847222, -782.13847222, -678.55347222, -539.25647222, -149.22347222,
    175.37552778, 463.88152778, 643.72952778, 803.40152778, -43.14647222, 36.03352778,
    -656.05547222, -435.77447222, -1112.94147222, 207.32652778, 523.49152778, 1186.61752778,
    2178.51052778, 2982.48652778, -17.41247222, -3017.44747222, -2213.43647222, -19.20647222};
static const double kAskapNorth[36] = {
    -105.22933333, -118.24033333, -97.73933333, -70.22133333, -73.72633333, 6.77066667,
    -215.20933333, -343.73933333, 5.32066667, 174.26666667, 235.30966667, 164.24666667,
    -469.23433333, -565.23833333, -263.22233333, 260.21066667, 417.28966667, 270.27966667,
    376.29066667, 209.79766667, 216.77666667, 230.75266667, -762.23633333, -1083.75233333,
    548.27766667, 562.27366667, 835.74566667, 1093.28166667, 647.98066667, 693.23866667,
    887.75566667, -2612.27533333, -2916.19433333, -2112.20733333, 887.76166667, 3084.83866667};

int tcv_current_uvw(float baseline, uint64_t *next, double *u, double *v, double *w)
{
    // current/tConvolveMPI/Benchmark.cc:168-258
    const double obslen = 12., scanlen = 5.;
    const int nScans = (int)(obslen * 3600. / scanlen);
    const double lat = -26.6970 * 3.141593 / 180.0;
    const double cdec = std::cos(lat), sdec = std::sin(lat);
    std::vector<double> BX, BY, BZ;
    double X[36], Y[36], Z[36];
    for (int i = 0; i < 36; i++) {
        X[i] = -kAskapNorth[i] * std::sin(lat);
        Y[i] = kAskapEast[i];
        Z[i] = kAskapNorth[i] * std::cos(lat);
    }
    for (int i = 0; i < 35; i++)
        for (int j = i + 1; j < 36; j++) {
            const double dX = X[i] - X[j], dY = Y[i] - Y[j], dZ = Z[i] - Z[j];
            if (dX * dX + dY * dY + dZ * dZ > (double)(baseline * baseline)) continue;
            BX.push_back(dX);
            BY.push_back(dY);
            BZ.push_back(dZ);
        }
    const int nBaselines = (int)BX.size();
    const int nSamples = nScans * nBaselines;
    if (!u || !v || !w || !next) return nSamples;
    const double maxint = 2147483647.0;
    for (int i = 0; i < nSamples; i++) {
        const int bl = (int)(nBaselines * (double(tcv_random_int(next)) / maxint));
        const double ha = obslen * 3.141593 / 12.0 * ((double(tcv_random_int(next)) / maxint) - 0.5);
        const double cha = std::cos(ha), sha = std::sin(ha);
        u[i] = sha * BX[bl] + cha * BY[bl];
        v[i] = -sdec * cha * BX[bl] + sdec * sha * BY[bl] + cdec * BZ[bl];
        w[i] = cdec * cha * BX[bl] - cdec * sha * BY[bl] + sdec * BZ[bl];
    }
    return nSamples;
}

tcv_status tcv_current_initC(double uvCellSize, int wSize, int support, int overSample,
                             double wCellSize, int *sSize, int *cOffset0, float *C,
                             long long *nTaps)
{
    // current/tConvolveMPI/Benchmark.cc:471-586
    if (!sSize || !cOffset0 || wSize < 1 || overSample < 1) return TCV_ERR_INVALID;
    if (wSize == 1) sSize[0] = 2 * support + 1;
    long long offsetCount = 0;
    std::vector<long long> off(wSize);
    for (int k = 0; k < wSize; k++) {
        const int wind = (int)(double)(k - wSize / 2);
        const double w = wind * wCellSize;
        cOffset0[k] = (int)offsetCount;
        off[k] = offsetCount;
        if (wSize > 1) {
            const float wPart = (float)(w / uvCellSize / uvCellSize);
            const float aPart = 7.0f;
            sSize[k] = (int)std::ceil(std::sqrt((double)(aPart * aPart + wPart * wPart)));
            sSize[k] += (sSize[k] + 1) % 2;
        }
        offsetCount += (long long)sSize[k] * sSize[k] * overSample * overSample;
    }
    if (nTaps) *nTaps = offsetCount;
    if (!C) return TCV_OK;
    std::complex<float> *Cv = reinterpret_cast<std::complex<float> *>(C);
    parallel_for(wSize, [&](int k) {
        const int wind = (int)(double)(k - wSize / 2);
        const double w = wind * wCellSize;
        const double fScale = wind != 0 ? uvCellSize * uvCellSize / w : 0.0;
        const int s = sSize[k];
        const int cCenter = s / 2;
        double sumC = 0.0;
        for (int osj = 0; osj < overSample; osj++)
            for (int osi = 0; osi < overSample; osi++) {
                const long long osOffset = (long long)s * s * (osi + overSample * osj) + off[k];
                for (int j = 0; j < s; j++) {
                    const double tj = double(j - cCenter) + double(osj) / double(overSample);
                    const double j2 = tj * tj;
                    for (int i = 0; i < s; i++) {
                        const long long cind = i + (long long)s * j + osOffset;
                        const double ti = double(i - cCenter) + double(osi) / double(overSample);
                        const double r2 = j2 + ti * ti;
                        float re = (float)std::exp(-r2), im = 0.0f;
                        if (wind != 0 && r2 < (double)(s / 2)) {
                            const float n = (float)std::sqrt(1. - r2 * fScale / w);
                            const float phase = (float)(-2. * 3.141593 * (r2 * fScale + w * ((double)n - 1.)));
                            const float a = (float)((double)n / w);
                            const float qr = std::sin(phase), qi = -std::cos(phase);
                            // Value(a) * Value(qr, qi), then C *= that (limited-range products)
                            const float tr = a * qr - 0.0f * qi, ti2 = a * qi + 0.0f * qr;
                            const float nr = re * tr - im * ti2, ni = re * ti2 + im * tr;
                            re = nr;
                            im = ni;
                        }
                        Cv[cind] = std::complex<float>(re, im);
                        sumC += (double)hypotf(re, im);
                    }
                }
            }
        const float normC = (float)(overSample * overSample / sumC);
        const long long n = (long long)s * s * overSample * overSample;
        for (long long i = 0; i < n; i++) {
            const float re = Cv[off[k] + i].real(), im = Cv[off[k] + i].imag();
            Cv[off[k] + i] = std::complex<float>(re * normC - im * 0.0f, re * 0.0f + im * normC);
        }
    });
    return TCV_OK;
}

}  // extern "C"
