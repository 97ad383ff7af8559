// Device FP64 Bessel functions of the first kind for the orders the
// transform supports (integer n <= 200, half-integer n + 1/2 <= 19/2).
// They replace GSL's gsl_sf_bessel_Jnu / gsl_sf_bessel_Jn that the
// reference calls per pair (src/expansions.c:71, :89, :161, :177, :285, :301).
//
// Methods (published): power series DLMF 10.2.2 where it does not cancel;
// Hankel's expansion DLMF 10.17.3 once it converges below 1e-17; Miller's
// backward recurrence (DLMF 3.6(vi)) normalised by J_0 + 2 sum J_2k = 1
// (DLMF 10.12.4) in between; half-integer orders via
// J_{n+1/2}(x) = sqrt(2x/pi) j_n(x) (DLMF 10.47.3) with the upward
// recurrence 10.51.1 (stable for n < x).  The argument is passed as an
// unevaluated sum hi + lo (FMA two-product of w and r) so the phase of the
// oscillatory branches is not rounded away at large w*r.
#pragma once
#include <hip/hip_runtime.h>

namespace nufht_dev {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kRt2 = 0.70710678118654752440084436210484904;

__device__ __forceinline__ void sincos_dd(double hi, double lo, double *s, double *c)
{
    double sh, ch;
    sincos(hi, &sh, &ch);
    *s = fma(lo, ch, sh);
    *c = fma(-lo, sh, ch);
}

// cos(q pi / 4), q in 0..7, exactly
__device__ __forceinline__ double cos_qpi4(int q)
{
    switch (q & 7) {
    case 0: return 1.0;
    case 1: case 7: return kRt2;
    case 2: case 6: return 0.0;
    case 3: case 5: return -kRt2;
    default: return -1.0;
    }
}

// Series for J_{n + half/2}(x), half in {0, 1}.
__device__ inline double series_j(int n, int half, double x)
{
    double h = 0.5 * x;
    double t0 = half ? sqrt(h) * (2.0 / sqrt(kPi)) : 1.0;
    for (int k = 1; k <= n; ++k) t0 *= h / ((double)k + (half ? 0.5 : 0.0));
    if (t0 == 0.0) return 0.0;
    double nu = (double)n + (half ? 0.5 : 0.0);
    double sum = t0, t = t0, h2 = h * h;
    for (int k = 1; k < 200; ++k) {
        t *= -h2 / ((double)k * ((double)k + nu));
        sum += t;
        if (fabs(t) < 1e-18 * fabs(sum)) break;
    }
    return sum;
}

// Hankel expansion; returns false if it does not reach 1e-17 relative.
__device__ inline bool hankel_j(double nu, double hi, double lo, double *out)
{
    double x = hi, mu4 = 4.0 * nu * nu;
    double P = 1.0, Q = 0.0, a = 1.0, prev = 1e300;
    bool conv = false;
    for (int k = 1; k < 120; ++k) {
        double odd = (double)(2 * k - 1);
        a *= (mu4 - odd * odd) / (8.0 * (double)k * x);
        double mag = fabs(a);
        if (mag > prev) break;
        prev = mag;
        if (k & 1) Q += ((((k - 1) >> 1) & 1) ? -a : a);
        else P += (((k >> 1) & 1) ? -a : a);
        if (mag < 1e-17 * (fabs(P) + fabs(Q)) || a == 0.0) { conv = true; break; }
    }
    if (!conv) return false;
    double s, c;
    sincos_dd(hi, lo, &s, &c);
    // rotation by (nu/2 + 1/4) pi: a multiple of pi/4 for our orders
    int q = (int)lround(2.0 * nu + 1.0) & 7;
    double cth = cos_qpi4(q), sth = cos_qpi4((q + 6) & 7);
    double cw = c * cth + s * sth;
    double sw = s * cth - c * sth;
    *out = sqrt(2.0 / (kPi * x)) * (P * cw - Q * sw);
    return true;
}

__device__ inline double miller_jn(int n, double x)
{
    double m = (double)(n > (int)x ? n : (int)x);
    int N = (int)(m + 20.0 + 4.0 * sqrt(m + 1.0) + 6.0 * cbrt(m + 1.0));
    N += N & 1;
    double jp1 = 0.0, j = 1e-280, saved = 0.0, sum = 0.0, tx = 2.0 / x;
    for (int k = N; k >= 1; --k) {
        double jm1 = (double)k * tx * j - jp1;
        jp1 = j;
        j = jm1;
        int km1 = k - 1;
        if (km1 == n) saved = j;
        if (km1 > 0 && !(km1 & 1)) sum += 2.0 * j;
        if (fabs(j) > 1e200) { j *= 1e-200; jp1 *= 1e-200; saved *= 1e-200; sum *= 1e-200; }
    }
    return saved / (sum + j);
}

__device__ inline double jn_int(int n, double hi, double lo)
{
    double x = hi;
    if (x == 0.0) return n == 0 ? 1.0 : 0.0;
    double h = 0.5 * x;
    if (h * h <= fmax(1.0, 0.5 * (double)(n + 1))) return series_j(n, 0, x);
    if (x >= 20.0) {
        double v;
        if (hankel_j((double)n, hi, lo, &v)) return v;
    }
    return miller_jn(n, x);
}

// spherical j_n(x), n <= 9
__device__ inline double sph_jn(int n, double hi, double lo)
{
    double x = hi;
    if (x == 0.0) return n == 0 ? 1.0 : 0.0;
    if (x <= fmax(2.0, (double)n)) return series_j(n, 1, x) * sqrt(kPi / (2.0 * x));
    double s, c;
    sincos_dd(hi, lo, &s, &c);
    double xi = 1.0 / x;
    double j0 = s * xi;
    if (n == 0) return j0;
    double j1 = (s * xi - c) * xi;
    for (int k = 1; k < n; ++k) {
        double j2 = (double)(2 * k + 1) * xi * j1 - j0;
        j0 = j1;
        j1 = j2;
    }
    return j1;
}

__device__ inline double jn_half(int n, double hi, double lo)
{
    double x = hi;
    if (x == 0.0) return 0.0;
    if (x <= fmax(2.0, (double)n)) return series_j(n, 1, x);
    return sqrt(2.0 * x / kPi) * sph_jn(n, hi, lo);
}

// J_{order + half/2}(w * r)
__device__ __forceinline__ double besselj_wr(int order, int half, double w, double r)
{
    double hi = w * r;
    double lo = fma(w, r, -hi);
    return half ? jn_half(order, hi, lo) : jn_int(order, hi, lo);
}

}  // namespace nufht_dev
