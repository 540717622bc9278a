/*
 * pg_oracle.c -- CPU restatement of PowerGrid's field-corrected SENSE path
 * (the oracle).  TEST INFRASTRUCTURE ONLY: see pg_oracle.h.
 *
 * Built by oracle/Makefile with gcc -O2 (no -ffast-math, no FMA contraction)
 * and OpenMP; orc_set_threads(1) gives the serial reference build.
 */
#define _GNU_SOURCE
#include "pg_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }
static inline int orc_nthreads(void) { return g_threads; }

/* griddingSupport.cpp:33-48 (Numerical Recipes in C, 2nd ed., bessi0) */
static double orc_bessi0(double x)
{
    double ax, ans, y;
    if ((ax = fabs(x)) < 3.75) {
        y = x / 3.75;
        y = y * y;
        ans = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
    } else {
        y = 3.75 / ax;
        ans = (exp(ax) / sqrt(ax)) * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
    }
    return ans;
}

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)

#define R float
#define SUF f
#define PI_R 3.1415926535897932384626433832795029f
#include "pg_oracle_impl.h"
#undef R
#undef SUF
#undef PI_R

#define R double
#define SUF d
#define PI_R 3.1415926535897932384626433832795029
#include "pg_oracle_impl.h"
#undef R
#undef SUF
#undef PI_R

/* Exact field-corrected NUDFT in double (double sin/cos):
 *   y_i = sum_j x_j exp(-i (2 pi k_i . r_j + fm_j t_i))     (Gdft.cpp:52-97)
 *   x_j = sum_i y_i exp(+i (2 pi k_i . r_j + fm_j t_i))     (Gdft.cpp:100-136) */
void orc_nudft_exact_fwd(double* out, const double* in, const double* kx, const double* ky,
                         const double* kz, const double* ix, const double* iy,
                         const double* iz, const double* fm, const double* t, int M, int N)
{
    const double tpi = 2.0 * M_PI;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < M; ++i) {
        double sr = 0, si = 0;
        for (int j = 0; j < N; ++j) {
            double ph = tpi * (kx[i] * ix[j] + ky[i] * iy[j] + kz[i] * iz[j]) + fm[j] * t[i];
            double c = cos(ph), s = sin(ph);
            sr += c * in[2 * j] + s * in[2 * j + 1];
            si += -s * in[2 * j] + c * in[2 * j + 1];
        }
        out[2 * i] = sr;
        out[2 * i + 1] = si;
    }
}

void orc_nudft_exact_adj(double* out, const double* in, const double* kx, const double* ky,
                         const double* kz, const double* ix, const double* iy,
                         const double* iz, const double* fm, const double* t, int M, int N)
{
    const double tpi = 2.0 * M_PI;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int j = 0; j < N; ++j) {
        double sr = 0, si = 0;
        for (int i = 0; i < M; ++i) {
            double ph = tpi * (kx[i] * ix[j] + ky[i] * iy[j] + kz[i] * iz[j]) + fm[j] * t[i];
            double c = cos(ph), s = sin(ph);
            sr += c * in[2 * i] - s * in[2 * i + 1];
            si += s * in[2 * i] + c * in[2 * i + 1];
        }
        out[2 * j] = sr;
        out[2 * j + 1] = si;
    }
}
