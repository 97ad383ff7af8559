/*
 * oracle/port.c -- TEST INFRASTRUCTURE (the oracle), never product code.
 *
 * CPU restatement of the reference's nufht_batch ("ref-algorithm CPU"):
 *   nufht_batch dispatch     src/fast_hankel_nufht.c:989-1107
 *   execute_boxes_batch      src/fast_hankel_nufht.c:923-987
 *   add_dir_batch            src/expansions.c:78-97
 *   add_loc_batch            src/expansions.c:212-337
 *   add_asy_batch            src/expansions.c:483-559
 * with the oracle's own J_nu (bessel.c) in place of GSL and the oracle's
 * type-3 NUFFT (nufft.c) or exact exponential sums in place of FINUFFT.
 * Deviation, on purpose: negative odd integer orders are negated once for
 * every batch size (the reference negates twice when batch == 1,
 * src/fast_hankel_nufht.c:916-920 and :1090-1095).
 * Columns are independent, so OpenMP splits the batch into one column
 * group per thread; with one thread the whole batch goes through each box
 * at once, as in the reference (single-threaded, FINUFFT nthreads = 1 at
 * src/expansions.c:347, :377).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Planted mistakes, for the tests that show the direct-sum comparison has
 * teeth (0 = none): 1 the odd (sine) Hankel terms' sign flipped, 2 the odd
 * terms' source weight r^(-2l-3/2) taken as r^(-2l-1/2), 3 the Wimp
 * expansion's alternating sign (-1)^(l-l0) dropped, 4 the Hankel phase
 * e^{i pi (-nu/2 - 1/4)} taken as e^{i pi (-nu/2 + 1/4)}. */
static int g_plant = 0;
void orc_set_plant(int which) { g_plant = which; }

static void add_dir_cols(double *gs, int ldg, double nu, const double *rs,
                         const double *cs, int ldc, const double *ws, int m,
                         int n, int B)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            double J = orc_besselj_wr(nu, ws[i], rs[j]);
            for (int t = 0; t < B; ++t) gs[i + (size_t)t * ldg] += cs[j + (size_t)t * ldc] * J;
        }
}

static void add_loc_cols(double *gs, int ldg, int nu, const double *rs,
                         const double *cs, int ldc, const double *ws, int m,
                         int n, int K, int B)
{
    int l0 = nu / 2, odd = nu & 1;
    double r_end = rs[m - 1];
    double *cheb = calloc((size_t)(K + 1) * B, sizeof(double));
    double *b1 = malloc(sizeof(double) * (K + 1));
    double *b2 = malloc(sizeof(double) * (l0 + K + odd + 1));
    for (int l = 0; l <= K; ++l) {
        double eps_l = (l == 0) ? 1.0 : 2.0;
        for (int k = 0; k < m; ++k) {
            double x = rs[k] / r_end;
            if (x > 1.0) x = 1.0;
            if (x < -1.0) x = -1.0;
            double th = acos(x), val;
            if (nu == 0) {
                double cl = cos((double)l * th);
                val = eps_l * (2.0 * cl * cl - 1.0);
            } else if (!odd) {
                val = eps_l * cos(2.0 * (double)l * th);
            } else {
                val = 2.0 * cos((double)(2 * l + 1) * th);
            }
            for (int t = 0; t < B; ++t) cheb[l + (size_t)t * (K + 1)] += val * cs[k + (size_t)t * ldc];
        }
    }
    int maxo = l0 + K + odd;
    for (int i = 0; i < n; ++i) {
        double z = ws[i] * r_end / 2.0;
        for (int k = 0; k <= maxo; ++k) b2[k] = orc_jn_int_dd(k, z, 0.0);
        for (int l = 0; l <= K; ++l) {
            int lo = l0 - l;
            if (lo < 0) lo = -lo;
            double v = b2[l0 + odd + l] * b2[lo];
            if (l > l0 && ((l - l0) & 1) && g_plant != 3) v = -v;
            b1[l] = v;
        }
        for (int t = 0; t < B; ++t) {
            double acc = 0.0;
            for (int l = 0; l <= K; ++l) acc += b1[l] * cheb[l + (size_t)t * (K + 1)];
            gs[i + (size_t)t * ldg] += acc;
        }
    }
    free(cheb); free(b1); free(b2);
}

static void add_asy_cols(double *gs, int ldg, double nu, const double *rs,
                         const double *cs, int ldc, const double *ws, int m,
                         int n, int K, double tol, const double *coef, int B,
                         int mode)
{
    double sq = sqrt(2.0 / M_PI);
    double pa = M_PI * (-nu / 2.0 + (g_plant == 4 ? 0.25 : -0.25));
    double complex ph = cos(pa) + I * sin(pa);
    double *cin = malloc(sizeof(double) * 2 * (size_t)m * B);
    double *cout = malloc(sizeof(double) * 2 * (size_t)n * B);
    for (int l = 0; l <= K; ++l) {
        for (int half = 0; half < 2; ++half) {
            double ex = -2.0 * l - 1.0 - half + 0.5;
            double exs = (g_plant == 2 && half) ? ex + 1.0 : ex;
            for (int t = 0; t < B; ++t)
                for (int k = 0; k < m; ++k) {
                    cin[2 * ((size_t)t * m + k)] = cs[k + (size_t)t * ldc] * pow(rs[k], exs);
                    cin[2 * ((size_t)t * m + k) + 1] = 0.0;
                }
            if (mode == 0) orc_nufft3_direct(m, rs, cin, +1, n, ws, cout, B);
            else orc_nufft3(m, rs, cin, +1, tol, n, ws, cout, B);
            double coeff = sq * ((l & 1) ? -1.0 : 1.0) * coef[2 * l + half];
            for (int t = 0; t < B; ++t)
                for (int i = 0; i < n; ++i) {
                    double complex v = (cout[2 * ((size_t)t * n + i)] +
                                        I * cout[2 * ((size_t)t * n + i) + 1]) * ph;
                    double sc = coeff * pow(ws[i], ex);
                    if (!half) gs[i + (size_t)t * ldg] += creal(v) * sc;
                    else gs[i + (size_t)t * ldg] += (g_plant == 1 ? 1.0 : -1.0) * cimag(v) * sc;
                }
        }
    }
    free(cin); free(cout);
}

int orc_nufht_batch(double nu, const double *rs, int m, const double *ws,
                    int n, const double *cs, int ldc, double *gs, int ldg,
                    int B, double tol, int asy_mode, int nthreads)
{
    if (!rs || !ws || m <= 0 || n <= 0) return -200;
    OrcParams p;
    int rc = orc_fill_params(&p, nu, tol, 0, 10000, NAN, 0, 0);
    if (rc) return rc;
    if (!cs || !gs || B <= 0) return -400;
    if (ldc < m || ldg < n) return -401;
    OrcBoxes bx;
    orc_generate_boxes(rs, m, ws, n, p.z_split, 0, 10000, p.K_loc < 0 ? 0 : p.K_loc, &bx);
    double a = fabs(nu);
    int nu_int = fabs(nu - nearbyint(nu)) < 1e-12;
    int use_direct = (long long)m * (long long)n < 10000LL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > B) nthreads = B;
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
    for (int grp = 0; grp < nthreads; ++grp) {
        int t0 = (int)((long long)B * grp / nthreads);
        int nb = (int)((long long)B * (grp + 1) / nthreads) - t0;
        if (nb <= 0) continue;
        double *g = gs + (size_t)t0 * ldg;
        const double *c = cs + (size_t)t0 * ldc;
        for (int t = 0; t < nb; ++t) memset(g + (size_t)t * ldg, 0, sizeof(double) * n);
        if (use_direct) {
            add_dir_cols(g, ldg, a, rs, c, ldc, ws, m, n, nb);
        } else if (!nu_int) {
            add_asy_cols(g, ldg, a, rs, c, ldc, ws, m, n, p.K_asy, tol, p.asy_coef, nb, asy_mode);
        } else {
            if (p.K_loc >= 0)
                for (int b = 0; b < bx.nloc; ++b) {
                    OrcBox q = bx.loc[b];
                    add_loc_cols(g + q.i0, ldg, (int)llround(a), rs + q.j0, c + q.j0, ldc,
                                 ws + q.i0, q.j1 - q.j0 + 1, q.i1 - q.i0 + 1, p.K_loc, nb);
                }
            for (int b = 0; b < bx.nasy; ++b) {
                OrcBox q = bx.asy[b];
                add_asy_cols(g + q.i0, ldg, a, rs + q.j0, c + q.j0, ldc, ws + q.i0,
                             q.j1 - q.j0 + 1, q.i1 - q.i0 + 1, p.K_asy, tol, p.asy_coef, nb,
                             asy_mode);
            }
            for (int b = 0; b < bx.ndir; ++b) {
                OrcBox q = bx.dir[b];
                add_dir_cols(g + q.i0, ldg, a, rs + q.j0, c + q.j0, ldc, ws + q.i0,
                             q.j1 - q.j0 + 1, q.i1 - q.i0 + 1, nb);
            }
        }
        if (nu < 0.0 && nu_int && (llround(nu) & 1))
            for (int t = 0; t < nb; ++t)
                for (int i = 0; i < n; ++i) g[i + (size_t)t * ldg] = -g[i + (size_t)t * ldg];
    }
    orc_boxes_free(&bx);
    return 0;
}
