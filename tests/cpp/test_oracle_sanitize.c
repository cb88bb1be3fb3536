/* Runs the oracle's gridder, degridder, offsets and Hogbom functions on small
 * random problems, including footprints on every grid edge, under
 * AddressSanitizer and UndefinedBehaviorSanitizer (host code only), and
 * checks two invariants: gridding is linear in the data, and degridding a grid
 * of ones gives each visibility the sum of its C slice. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "tcv_oracle.h"

#define CHECK(c)                                              \
    do {                                                      \
        if (!(c)) {                                           \
            printf("FAIL line %d: %s\n", __LINE__, #c);       \
            return 1;                                         \
        }                                                     \
    } while (0)

static uint64_t st = 88172645463325252ull;
static double urand(void)
{
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
}

int main(void)
{
    const int gSize = 96, S = 17, half = S / 2, os = 4, planes = 3, nVis = 400;
    const long long slice = (long long)S * S, nSlices = (long long)os * os * planes;
    float *C = malloc(sizeof(float) * 2 * slice * nSlices);
    for (long long i = 0; i < 2 * slice * nSlices; i++) C[i] = (float)(urand() - 0.5);
    int *iu = malloc(sizeof(int) * nVis), *iv = malloc(sizeof(int) * nVis);
    int *co = malloc(sizeof(int) * nVis);
    float *data = malloc(sizeof(float) * 2 * nVis), *data2 = malloc(sizeof(float) * 2 * nVis);
    for (int k = 0; k < nVis; k++) {
        /* the first visibilities sit on the four edges of the valid range */
        const int edge = k % 8;
        iu[k] = edge == 0 ? half : edge == 1 ? gSize - half - 1 : half + (int)(urand() * (gSize - S));
        iv[k] = edge == 2 ? 0 : edge == 3 ? gSize - S : (int)(urand() * (gSize - S + 1));
        co[k] = (int)((long long)(urand() * nSlices) * slice);
        data[2 * k] = (float)(urand() - 0.5);
        data[2 * k + 1] = (float)(urand() - 0.5);
        data2[2 * k] = 2.0f * data[2 * k];
        data2[2 * k + 1] = 2.0f * data[2 * k + 1];
    }
    const size_t gBytes = sizeof(float) * 2 * (size_t)gSize * gSize;
    float *g1 = calloc(1, gBytes), *g2 = calloc(1, gBytes);
    orc_grid(nVis, data, iu, iv, co, NULL, NULL, S, C, g1, gSize);
    orc_grid(nVis, data2, iu, iv, co, NULL, NULL, S, C, g2, gSize);
    double worst = 0, mx = 0;
    for (long long i = 0; i < 2LL * gSize * gSize; i++) {
        worst = fmax(worst, fabs(2.0 * g1[i] - g2[i]));
        mx = fmax(mx, fabs(g2[i]));
    }
    CHECK(mx > 0 && worst <= 1e-5 * mx);

    float *ones = malloc(gBytes), *out = malloc(sizeof(float) * 2 * nVis);
    for (long long i = 0; i < (long long)gSize * gSize; i++) {
        ones[2 * i] = 1.0f;
        ones[2 * i + 1] = 0.0f;
    }
    orc_degrid(nVis, ones, gSize, iu, iv, co, NULL, NULL, S, C, out);
    for (int k = 0; k < nVis; k++) {
        double sr = 0, si = 0;
        for (long long t = 0; t < slice; t++) {
            sr += C[2 * (co[k] + t)];
            si += C[2 * (co[k] + t) + 1];
        }
        CHECK(fabs(sr - out[2 * k]) <= 1e-4 && fabs(si - out[2 * k + 1]) <= 1e-4);
    }

    /* attic offsets on a few thousand LCG samples */
    const int nS = 2000;
    double *u = malloc(sizeof(double) * nS), *v = malloc(sizeof(double) * nS);
    double *w = malloc(sizeof(double) * nS), freq[1];
    uint64_t next = 1;
    orc_attic_uvw(nS, 2000, &next, u, v, w);
    orc_attic_freq(1, freq);
    int support;
    double wcs;
    orc_attic_initC(freq[0], 5.0, 33, 2000, 8, &support, &wcs, NULL);
    CHECK(support == 64);
    int *oiu = malloc(sizeof(int) * nS), *oiv = malloc(sizeof(int) * nS);
    int *oco = malloc(sizeof(int) * nS);
    orc_attic_initCOffset(nS, u, v, w, 1, freq, 5.0, wcs, 33, 4096, support, 8, oiu, oiv, oco);
    for (int k = 0; k < nS; k++) CHECK(oiu[k] >= 0 && oiu[k] < 4096 && oiv[k] >= 0 && oco[k] >= 0);

    /* Hogbom on a small image with a PSF whose peak is off the diagonal */
    const int W = 40, PW = 24;
    float *res = malloc(sizeof(float) * 2 * W * W), *psf = malloc(sizeof(float) * 2 * PW * PW);
    float *model = calloc(2 * W * W, sizeof(float));
    for (int i = 0; i < 2 * W * W; i++) res[i] = (float)(urand() - 0.5);
    for (int i = 0; i < 2 * PW * PW; i++) psf[i] = (float)(0.1 * (urand() - 0.5));
    psf[2 * (3 * PW + 17)] = 4.0f;
    float pv, dv;
    long long pp, dp;
    orc_deconvolve(res, W, psf, PW, model, 50, 0.1f, &pv, &pp, &dv, &dp);
    CHECK(pp == 3 * PW + 17 && pv == 4.0f);

    free(C); free(iu); free(iv); free(co); free(data); free(data2); free(g1); free(g2);
    free(ones); free(out); free(u); free(v); free(w); free(oiu); free(oiv); free(oco);
    free(res); free(psf); free(model);
    printf("oracle sanitize ok\n");
    return 0;
}
