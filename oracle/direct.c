/*
 * oracle/direct.c -- TEST INFRASTRUCTURE (the oracle), never product code.
 *
 * FP64 direct sum g[i,t] = sum_j c[j,t] J_nu(w_i r_j): the literal
 * definition the reference evaluates in add_dir_batch
 * (src/expansions.c:78-97), here with Neumaier-compensated accumulation and
 * the oracle's own J_nu.  Order conventions follow the reference's result
 * for valid inputs: negative integer nu gives (-1)^nu J_|nu|
 * (src/fast_hankel_nufht.c:1090-1095 applied once), negative half-integer
 * nu gives J_|nu| (src/fast_hankel_nufht.c:932, 1054).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

static double order_sign(double nu)
{
    if (nu < 0.0 && fabs(nu - nearbyint(nu)) < 1e-12 && (llround(nu) & 1)) return -1.0;
    return 1.0;
}

static void one_row(double nu, const double *r, int m, const double *c,
                    int ldc, int B, double w, double *out, int ldo)
{
    double a = fabs(nu);
    double sgn = order_sign(nu);
    double sum[64], comp[64];
    for (int t0 = 0; t0 < B; t0 += 64) {
        int nb = B - t0 < 64 ? B - t0 : 64;
        for (int t = 0; t < nb; ++t) { sum[t] = 0.0; comp[t] = 0.0; }
        for (int j = 0; j < m; ++j) {
            double J = orc_besselj_wr(a, w, r[j]);
            for (int t = 0; t < nb; ++t) {
                double v = c[j + (size_t)(t0 + t) * ldc] * J;
                double s = sum[t] + v;
                if (fabs(sum[t]) >= fabs(v)) comp[t] += (sum[t] - s) + v;
                else comp[t] += (v - s) + sum[t];
                sum[t] = s;
            }
        }
        for (int t = 0; t < nb; ++t) out[(size_t)(t0 + t) * ldo] = sgn * (sum[t] + comp[t]);
    }
}

void orc_direct(double nu, const double *r, int m, const double *c, int ldc,
                int B, const double *w, int n, double *g, int ldg, int nthreads)
{
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < n; ++i) one_row(nu, r, m, c, ldc, B, w[i], g + i, ldg);
}

void orc_direct_rows(double nu, const double *r, int m, const double *c,
                     int ldc, int B, const double *w, const int *rows, int nr,
                     double *out, int nthreads)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int k = 0; k < nr; ++k) one_row(nu, r, m, c, ldc, B, w[rows[k]], out + k, nr);
}
