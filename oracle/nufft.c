/*
 * oracle/nufft.c -- TEST INFRASTRUCTURE (the oracle), never product code.
 *
 * CPU 1-D type-3 NUFFT, f_k = sum_j c_j exp(i*iflag*x_j*s_k).  The
 * reference obtains this from FINUFFT's finufft1d3 / finufft1d3many
 * (src/expansions.c:339-397; FINUFFT is not vendored and its version is
 * not pinned: Makefile:27-28, python/setup.py:29).  FINUFFT is absent
 * here, so the published method is restated:
 *   - Lee & Greengard, SIAM J. Sci. Comput. 2005 (type 3 as spread +
 *     type-2 transform + deconvolution at the nonuniform frequencies);
 *   - Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput. 2019 (the
 *     "exponential of semicircle" kernel exp(beta(sqrt(1-u^2)-1)),
 *     upsampling 2, width w = ceil(log10(1/eps)) + 2, beta = 2.30 w).
 * The uniform FFT is a plain iterative radix-2 transform on power-of-two
 * grids; kernel weights and phases are computed once per point and shared
 * by all transforms of a batch.  Pinned against orc_nufft3_direct
 * (brute-force sums).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void orc_nufft3_direct(int m, const double *x, const double *c, int iflag,
                       int n, const double *s, double *f, int ntr)
{
    for (int t = 0; t < ntr; ++t)
        for (int k = 0; k < n; ++k) {
            double re = 0.0, im = 0.0;
            for (int j = 0; j < m; ++j) {
                double hi = x[j] * s[k], lo = fma(x[j], s[k], -hi);
                double sn = sin(hi) + lo * cos(hi), cs = cos(hi) - lo * sin(hi);
                sn *= (double)iflag;
                double cr = c[2 * ((size_t)t * m + j)], ci = c[2 * ((size_t)t * m + j) + 1];
                re += cr * cs - ci * sn;
                im += cr * sn + ci * cs;
            }
            f[2 * ((size_t)t * n + k)] = re;
            f[2 * ((size_t)t * n + k) + 1] = im;
        }
}

static void fft_pow2(double complex *a, int N, int sign)
{
    for (int i = 1, j = 0; i < N; ++i) {
        int bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { double complex t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    /* twiddles e^{sign 2 pi i k / N}, k < N/2, computed once */
    double complex *tw = malloc(sizeof(double complex) * (size_t)(N / 2 + 1));
    for (int k = 0; k < N / 2; ++k) {
        double ang = sign * 2.0 * M_PI * (double)k / (double)N;
        tw[k] = cos(ang) + I * sin(ang);
    }
    for (int len = 2; len <= N; len <<= 1) {
        int step = N / len;
        for (int i = 0; i < N; i += len)
            for (int k = 0; k < len / 2; ++k) {
                double complex u = a[i + k], v = a[i + k + len / 2] * tw[k * step];
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
    free(tw);
}

/* Gauss-Legendre nodes/weights on [-1,1] by Newton on P_q. */
static void gauss_legendre(int q, double *xs, double *ws)
{
    for (int i = 0; i < q; ++i) {
        double x = cos(M_PI * (i + 0.75) / (q + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= q; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            double dp = q * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= q; ++k) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1; p1 = p2;
        }
        double dp = q * (x * p1 - p0) / (x * x - 1.0);
        xs[i] = x;
        ws[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

static double es(double u, double beta)
{
    double t = 1.0 - u * u;
    return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

/* Fourier transform of the kernel with half-width a: int psi(x) cos(kx) dx */
static double es_ft(double k, double a, double beta, int q, const double *gx,
                    const double *gw)
{
    double s = 0.0;
    for (int i = 0; i < q; ++i) s += gw[i] * es(gx[i], beta) * cos(k * a * gx[i]);
    return s * a;
}

/* F(k) = int_{-1}^{1} es(u) cos(k u) du as an even Chebyshev series on
 * [-kmax, kmax] (in y = 2 (k/kmax)^2 - 1), so the per-target deconvolution
 * costs one short Clenshaw sum instead of a quadrature. */
#define ORC_NFT 48
typedef struct { double kmax, c[ORC_NFT]; } FtTable;

static void ft_table(FtTable *t, double kmax, double beta, int q, const double *gx,
                     const double *gw)
{
    enum { Q = 128 };
    double F[Q];
    t->kmax = kmax;
    for (int i = 0; i < Q; ++i) {
        double y = cos(M_PI * (i + 0.5) / Q);
        F[i] = es_ft(kmax * sqrt(fmax(0.0, 0.5 * (y + 1.0))), 1.0, beta, q, gx, gw);
    }
    for (int n = 0; n < ORC_NFT; ++n) {
        double acc = 0.0;
        for (int i = 0; i < Q; ++i) acc += F[i] * cos(n * M_PI * (i + 0.5) / Q);
        t->c[n] = acc * (n ? 2.0 : 1.0) / Q;
    }
}

static double ft_eval(const FtTable *t, double k)
{
    double x = k / t->kmax, y = 2.0 * x * x - 1.0, b1 = 0.0, b2 = 0.0;
    for (int n = ORC_NFT - 1; n >= 1; --n) {
        double b0 = 2.0 * y * b1 - b2 + t->c[n];
        b2 = b1;
        b1 = b0;
    }
    return y * b1 - b2 + t->c[0];
}

int orc_nufft3(int m, const double *x, const double *c, int iflag, double eps,
               int n, const double *s, double *f, int ntr)
{
    if (m <= 0 || n <= 0) return 0;
    int w = (int)ceil(log10(1.0 / eps)) + 2;
    if (w < 4) w = 4;
    if (w > 16) w = 16;
    double beta = 2.30 * w, sigma = 2.0;
    double xmin = x[0], xmax = x[0], smin = s[0], smax = s[0];
    for (int j = 1; j < m; ++j) { xmin = fmin(xmin, x[j]); xmax = fmax(xmax, x[j]); }
    for (int k = 1; k < n; ++k) { smin = fmin(smin, s[k]); smax = fmax(smax, s[k]); }
    double C = 0.5 * (xmin + xmax), X = 0.5 * (xmax - xmin);
    double D = 0.5 * (smin + smax), S = 0.5 * (smax - smin);
    /* keep the grid meaningful for degenerate (single-point) ranges */
    if (S < 1e-12 * (fabs(D) + 1.0)) S = 1e-12 * (fabs(D) + 1.0);
    double h = M_PI / (sigma * S);
    int L = (int)ceil(X / h) + w / 2 + 2;
    int Ng = 2 * L + 1;
    int Nf = 1;
    while (Nf < (int)(sigma * Ng) || Nf < 2 * w + 4) Nf <<= 1;
    double delta = 2.0 * M_PI / ((double)Nf * h);
    double ah = 0.5 * w * h, ad = 0.5 * w * delta;
    int q = 200;
    double *gx = malloc(sizeof(double) * q), *gw = malloc(sizeof(double) * q);
    gauss_legendre(q, gx, gw);
    FtTable ft;
    ft_table(&ft, 0.5 * M_PI * w, beta, q, gx, gw);
    double *dec = malloc(sizeof(double) * Ng);
    for (int l = -L; l <= L; ++l) dec[l + L] = 1.0 / (ad * ft_eval(&ft, (double)l * h * ad));
    /* spreading weights are shared by all transforms of the batch */
    double *tw = malloc(sizeof(double) * (size_t)m * (w + 2));
    int *tl0 = malloc(sizeof(int) * (size_t)m);
    double complex *ph = malloc(sizeof(double complex) * (size_t)m);
    for (int j = 0; j < m; ++j) {
        double xt = x[j] - C;
        int l0 = (int)ceil((xt - ah) / h);
        tl0[j] = l0;
        for (int k = 0; k < w + 2; ++k) tw[(size_t)j * (w + 2) + k] = es(((double)(l0 + k) * h - xt) / ah, beta);
        ph[j] = cos(iflag * D * xt) + I * sin(iflag * D * xt);
    }
    double *iw = malloc(sizeof(double) * (size_t)n * (w + 2));
    int *im0 = malloc(sizeof(int) * (size_t)n);
    double complex *post = malloc(sizeof(double complex) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double st = s[k] - D;
        int m0 = (int)ceil((st - ad) / delta);
        im0[k] = m0;
        for (int q2 = 0; q2 < w + 2; ++q2) iw[(size_t)k * (w + 2) + q2] = es((st - (double)(m0 + q2) * delta) / ad, beta);
        double phk = iflag * s[k] * C;
        post[k] = (cos(phk) + I * sin(phk)) * h * delta / (ah * ft_eval(&ft, st * ah));
    }
    double complex *grid = malloc(sizeof(double complex) * Nf);
    double complex *H = malloc(sizeof(double complex) * Ng);
    for (int t = 0; t < ntr; ++t) {
        memset(H, 0, sizeof(double complex) * Ng);
        for (int j = 0; j < m; ++j) {
            double complex v = (c[2 * ((size_t)t * m + j)] + I * c[2 * ((size_t)t * m + j) + 1]) * ph[j];
            const double *wt = tw + (size_t)j * (w + 2);
            for (int k = 0; k < w + 2; ++k) {
                int l = tl0[j] + k;
                if (l >= -L && l <= L) H[l + L] += v * wt[k];
            }
        }
        memset(grid, 0, sizeof(double complex) * Nf);
        for (int l = -L; l <= L; ++l) grid[((l % Nf) + Nf) % Nf] = H[l + L] * dec[l + L];
        fft_pow2(grid, Nf, iflag);
        for (int k = 0; k < n; ++k) {
            double complex acc = 0.0;
            const double *wt = iw + (size_t)k * (w + 2);
            for (int q2 = 0; q2 < w + 2; ++q2) {
                int mm = im0[k] + q2;
                acc += grid[((mm % Nf) + Nf) % Nf] * wt[q2];
            }
            acc *= post[k];
            f[2 * ((size_t)t * n + k)] = creal(acc);
            f[2 * ((size_t)t * n + k) + 1] = cimag(acc);
        }
    }
    free(gx); free(gw); free(dec); free(grid); free(H);
    free(tw); free(tl0); free(ph); free(iw); free(im0); free(post);
    return 0;
}
