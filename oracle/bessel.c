/*
 * oracle/bessel.c -- TEST INFRASTRUCTURE (the oracle), never product code.
 *
 * FP64 Bessel functions of the first kind J_nu(x) for the orders the
 * reference accepts (integer |nu| <= 200, half-integer |nu| <= 19/2,
 * src/fast_hankel_nufht.c:336-344).  They stand in for GSL's
 * gsl_sf_bessel_Jnu / gsl_sf_bessel_Jn, which the reference calls at
 * src/expansions.c:71, :89, :161, :177, :285, :301 (GSL is not installed,
 * so its arithmetic is restated from the published methods below).
 *
 * Methods (published, not taken from any library source):
 *   - power series, DLMF 10.2.2, where the terms do not cancel;
 *   - Hankel's asymptotic expansion, DLMF 10.17.3, with the coefficients
 *     a_k(nu) of DLMF 10.17.1, used only when the series converges to
 *     below 1e-17 relative;
 *   - Miller's backward recurrence (DLMF 3.6(vi)) normalised with
 *     J_0 + 2 sum_k J_2k = 1 (DLMF 10.12.4) otherwise;
 *   - half-integer orders through J_{n+1/2}(x) = sqrt(2x/pi) j_n(x)
 *     (DLMF 10.47.3) and the upward recurrence 10.51.1, which is stable
 *     for n < x.
 * The argument x = w*r is carried as an unevaluated sum hi + lo (an FMA
 * two-product), so the oscillatory phase of J_nu(w r) is resolved to far
 * below one ulp of x even at x ~ 1e6.
 *
 * Parity pins: tests/golden/bessel_golden.npz (mpmath 1.3.0, 30 digits),
 * the three-term recurrence, and the closed forms of J_{1/2}, J_{3/2}.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* sin/cos of (hi + lo) with |lo| <= ulp(hi): first-order correction. */
static void sincos_dd(double hi, double lo, double *s, double *c)
{
    double sh = sin(hi), ch = cos(hi);
    *s = sh + lo * ch;
    *c = ch - lo * sh;
}

/* Power series DLMF 10.2.2 for J_nu, nu = n + half (half in {0, 0.5}). */
static double series_j(int n, int half, double x)
{
    double h = 0.5 * x;
    double t0;
    if (!half) {
        t0 = 1.0;
        for (int k = 1; k <= n; ++k) t0 *= h / (double)k;
    } else {
        /* (x/2)^(n+1/2) / Gamma(n+3/2), Gamma(3/2) = sqrt(pi)/2 */
        t0 = sqrt(h) / (0.5 * sqrt(ORC_PI));
        for (int k = 1; k <= n; ++k) t0 *= h / ((double)k + 0.5);
    }
    if (t0 == 0.0) return 0.0;
    double nu = (double)n + (half ? 0.5 : 0.0);
    double sum = t0, t = t0, h2 = h * h;
    for (int k = 1; k < 400; ++k) {
        t *= -h2 / ((double)k * ((double)k + nu));
        sum += t;
        if (fabs(t) < 1e-18 * fabs(sum)) break;
    }
    return sum;
}

/* Hankel expansion DLMF 10.17.3.  Returns 1 and the value when the series
 * reaches 1e-17 relative before its terms start to grow, else 0. */
static int hankel_j(double nu, double hi, double lo, double *out)
{
    double x = hi;
    double mu4 = 4.0 * nu * nu;
    double P = 1.0, Q = 0.0;
    double a = 1.0;          /* a_k(nu) / x^k */
    double prev = INFINITY;
    int conv = 0;
    for (int k = 1; k < 200; ++k) {
        double odd = (double)(2 * k - 1);
        a *= (mu4 - odd * odd) / (8.0 * (double)k * x);
        double mag = fabs(a);
        if (mag > prev && mag > 1e-300) break;      /* diverging: give up */
        prev = mag;
        /* k even -> P with sign (-1)^(k/2); k odd -> Q with sign (-1)^((k-1)/2) */
        if (k & 1) Q += ((((k - 1) / 2) & 1) ? -a : a);
        else       P += (((k / 2) & 1) ? -a : a);
        if (mag < 1e-17 * (fabs(P) + fabs(Q))) { conv = 1; break; }
        if (a == 0.0) { conv = 1; break; }
    }
    if (!conv) return 0;
    /* omega = x - (nu/2 + 1/4) pi;  cos/sin of the shift are exact. */
    double s, c;
    sincos_dd(hi, lo, &s, &c);
    double th = 0.5 * nu + 0.25;          /* multiples of pi/4 for our orders */
    double cth = cos(th * ORC_PI), sth = sin(th * ORC_PI);
    /* snap the constant rotation to its exact value (multiples of pi/4) */
    double q = th * 4.0;
    long qi = lround(q);
    if (fabs(q - (double)qi) < 1e-12) {
        static const double ct[8] = {1, 0.70710678118654752440084436210484904, 0,
                                     -0.70710678118654752440084436210484904, -1,
                                     -0.70710678118654752440084436210484904, 0,
                                     0.70710678118654752440084436210484904};
        long m = ((qi % 8) + 8) % 8;
        cth = ct[m];
        sth = ct[(m + 6) % 8];            /* sin(t) = cos(t - pi/2) */
    }
    double cw = c * cth + s * sth;
    double sw = s * cth - c * sth;
    *out = sqrt(2.0 / (ORC_PI * x)) * (P * cw - Q * sw);
    return 1;
}

/* Miller backward recurrence for integer order n, DLMF 3.6(vi). */
static double miller_jn(int n, double x)
{
    double m = (double)(n > x ? n : (int)x);
    int N = (int)(m + 20.0 + 4.0 * sqrt(m + 1.0) + 2.0 * cbrt(m + 1.0) * 3.0);
    if (N & 1) N++;
    double jp1 = 0.0, j = 1e-280, saved = 0.0, sum = 0.0;
    double two_over_x = 2.0 / x;
    for (int k = N; k >= 1; --k) {
        double jm1 = (double)k * two_over_x * j - jp1;   /* J_{k-1} */
        jp1 = j;
        j = jm1;
        int km1 = k - 1;
        if (km1 == n) saved = j;
        if (km1 > 0 && !(km1 & 1)) sum += 2.0 * j;
        if (fabs(j) > 1e200) {
            j *= 1e-200; jp1 *= 1e-200; saved *= 1e-200; sum *= 1e-200;
        }
    }
    sum += j;                              /* J_0 */
    return saved / sum;
}

double orc_jn_int_dd(int n, double hi, double lo)
{
    double x = hi;
    if (x == 0.0) return n == 0 ? 1.0 : 0.0;
    double h = 0.5 * x;
    if (h * h <= fmax(1.0, 0.5 * (double)(n + 1))) return series_j(n, 0, x);
    if (x >= 20.0) {
        double v;
        if (hankel_j((double)n, hi, lo, &v)) return v;
    }
    /* Below the Hankel range x is modest, so the rounding of w*r (lo) moves
     * J by at most |lo| ~ x * 1.1e-16 and is ignored. */
    return miller_jn(n, x);
}

double orc_jn_half_dd(int n, double hi, double lo)
{
    double x = hi;
    if (x == 0.0) return 0.0;
    if (x <= fmax(2.0, (double)n)) {
        double v = series_j(n, 1, x);
        if (lo != 0.0) {
            /* J'_nu = (nu/x) J_nu - J_{nu+1} */
            double nu = (double)n + 0.5;
            v += lo * (nu / x * v - series_j(n + 1, 1, x));
        }
        return v;
    }
    double s, c;
    sincos_dd(hi, lo, &s, &c);
    double xi = 1.0 / x;
    double j0 = s * xi;
    if (n == 0) return sqrt(2.0 * x / ORC_PI) * j0;
    double j1 = s * xi * xi - c * xi;
    for (int k = 1; k < n; ++k) {
        double j2 = (double)(2 * k + 1) * xi * j1 - j0;
        j0 = j1;
        j1 = j2;
    }
    return sqrt(2.0 * x / ORC_PI) * j1;
}

/* J_nu(w*r) for nu >= 0 integer or half-integer; the product is formed as
 * an exact two-product so the phase is not rounded away. */
double orc_besselj_wr(double nu, double w, double r)
{
    double hi = w * r;
    double lo = fma(w, r, -hi);
    double a = fabs(nu);
    long twice = lround(2.0 * a);
    if (twice & 1) return orc_jn_half_dd((int)((twice - 1) / 2), hi, lo);
    return orc_jn_int_dd((int)(twice / 2), hi, lo);
}

double orc_besselj(double nu, double x)
{
    return orc_besselj_wr(nu, x, 1.0);
}
