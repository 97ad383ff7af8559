/*
 * oracle/tables.c -- TEST INFRASTRUCTURE (the oracle), never product code.
 *
 * Restates the reference's plan-table generator, src/bounds.c:63-324, step
 * for step (same operation order, so glibc reproduces the same doubles the
 * reference would compute on this machine):
 *   a_coef            src/bounds.c:63-82   (DLMF 10.17.1 coefficients)
 *   asy_error_bound   src/bounds.c:120-131
 *   newton_asy        src/bounds.c:141-178 (x0 = 1, rel 1e-8, central diff.)
 *   psi / wimp bound  src/bounds.c:223-250
 *   table fills       src/bounds.c:93-112, :194-214, :265-297
 * Pinned against the reference's own data/nufht_tables.cache (fixture slice
 * tests/golden/ref_tables_slice.npz) and against oracle/_ref when built.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

static const double TOLS[12] = {1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9,
                                1e-10, 1e-11, 1e-12, 1e-13, 1e-14, 1e-15};

double orc_a_coef(int k, double nu)
{
    if (k == 0) return 1.0;
    double prod = 1.0;
    for (int j = 1; j <= 2 * k - 1; j += 2) prod *= 4.0 * nu * nu - (double)(j * j);
    double fact = 1.0;
    for (int j = 2; j <= k; ++j) fact *= (double)j;
    return prod / (fact * pow(8.0, (double)k));
}

double orc_asy_error_bound(double nu, int K, double z)
{
    double a1 = fabs(orc_a_coef(2 * K, nu));
    double a2 = fabs(orc_a_coef(2 * K + 1, nu));
    double z1 = pow(z, (double)(2 * K));
    double z2 = pow(z, (double)(2 * K + 1));
    double pref = sqrt(2.0 / (M_PI * z));
    return pref * (a1 / z1 + a2 / z2);
}

static double newton_asy(double nu, int K, double tol)
{
    double x = 1.0;
    for (int it = 0; it < 1000; ++it) {
        double f = orc_asy_error_bound(nu, K, x) - tol;
        double h = 1e-5 * (fabs(x) > 1.0 ? fabs(x) : 1.0);
        double fp = orc_asy_error_bound(nu, K, x + h) - tol;
        double fm = orc_asy_error_bound(nu, K, x - h) - tol;
        double df = (fp - fm) / (2.0 * h);
        if (fabs(df) < 1e-16) break;
        double dx = -f / df;
        x += dx;
        if (x <= 0.0) x = fabs(x);
        if (fabs(dx / x) <= 1e-8) break;
    }
    return x;
}

static double psi_p(double p)
{
    double r = sqrt(fmax(0.0, 1.0 - p * p));
    return log(p) + r - log(1.0 + r);
}

double orc_wimp_error_bound(double nu, int K, double z)
{
    double bK = psi_p(z / (2.0 * K + 2.0 + nu));
    double tkn = 2.0 * K - nu;
    if (tkn > 0.0 && z / tkn <= 1.0) {
        double cK = psi_p(z / (2.0 * K + 2.0 - nu));
        double num = 2.0 * exp(bK * (nu / 2.0 + K + 1.0) + cK * (-nu / 2.0 + K + 1.0));
        return num / (1.0 - exp(bK + cK));
    }
    double num = 2.0 * exp(bK * (nu / 2.0 + K + 1.0));
    return num / (1.0 - exp(bK));
}

void orc_generate_tables(double *asy_coef, double *asy_z, int *wimp_k)
{
    memset(asy_coef, 0, sizeof(double) * ORC_NU_ROWS * ORC_ASY_COEF);
    memset(asy_z, 0, sizeof(double) * ORC_TI * ORC_TJ * ORC_TK);
    memset(wimp_k, 0, sizeof(int) * ORC_TI * ORC_TJ * ORC_TK);
    for (int i = 0; i <= 400; ++i)
        for (int J = 0; J <= 21; ++J)
            asy_coef[i * ORC_ASY_COEF + J] = orc_a_coef(J, 0.5 * (double)i);
    for (int nu = 0; nu <= 200; ++nu)
        for (int j = 0; j < 12; ++j)
            for (int K = 1; K <= 10; ++K) {
                size_t at = ((size_t)nu * ORC_TJ + j) * ORC_TK + (K - 1);
                double z = newton_asy((double)nu, K, TOLS[j]);
                asy_z[at] = z;
                int wk = (int)ceil((z - (double)nu) / 2.0);
                if (wk < 1) wk = 1;
                while (!(orc_wimp_error_bound((double)nu, wk, z) < TOLS[j])) wk++;
                wimp_k[at] = wk;
            }
}
