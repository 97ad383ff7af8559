// Device kernels of the MI355X NUFHT engine.
//
// Layout in HBM (all column-major, as in the reference ABI):
//   r[m], w[n]            ascending grids (FP64)
//   cs[t*ld_cs + j]       coefficients, gs[t*ld_gs + i] outputs (FP64)
//   coef[box][l][t]       local-expansion moments (FP64)
//   grid[box][e*B+t][Nf]  asymptotic-stage uniform grids, complex (FP64 or
//                         FP32), one length-Nf row per (exponent e, column t)
//
// Every stage is one launch over a flat list of work items covering all
// boxes of all transforms in the program; items carry their box, index
// range and column chunk.  Contributions to gs are accumulated with FP64
// global atomics (boxes of one transform share output rows).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>

#include "bessel_dev.h"

namespace nufht_dev {

constexpr int kMaxA = 24;

struct TfDesc {
    const double *r, *w;
    const double *cs;
    double *gs;
    long long ld_cs, ld_gs;
    int m, n, B;
    int order, half;       // |nu| = order + half/2
    int K_asy, K_loc;
    int kw;                // ES kernel width in grid points
    int ft_off;            // offset of the kernel-transform table in ft[]
    int cheb_off;          // half-integer local expansion matrix in cheb[]
    int jtab_n;            // unit intervals of the direct-sum Bessel table (0: none)
    long long jtab_off;    // offset of that table in jtab[]
    double beta;           // ES shape parameter
    double kmax;           // domain of the kernel-transform table
    double sign;           // -1 for negative odd integer orders
    double a[kMaxA];       // Hankel coefficients a_k(|nu|)
};

// rows [i0, i0 + ni) of transform tf, columns [col0, col0 + 16), sources in
// the ranges rng[roff .. roff + rcnt) ({j0, count} each)
struct DirItem { int tf, i0, ni, roff, rcnt, tpr, col0, pad; };
struct LocBoxD { int tf, i0, ni, j0, nj, coef_off; double r_end; };
struct LocItem { int box, start, count, col0; };
// targets [i0, i0 + count) of transform tf, columns [col0, col0 + BMAX), the
// local boxes boxlist[roff .. roff + rcnt) that cover them
struct LocEvalItem { int tf, i0, count, roff, rcnt, col0; };
struct AsyBoxD {
    int tf, i0, ni, j0, nj, E, L, Nf;
    long long goff;
    double C, D, h, delta, a1, a2, rho;
};
struct AsyItem { int box, start, count, col0, tb, pad; };
struct ZeroItem { int tf, col, start, count; };

constexpr int kNFT = 40;   // Chebyshev terms of the kernel transform table

// F(k) = int_{-1}^{1} exp(beta (sqrt(1-u^2) - 1)) cos(k u) du, from its even
// Chebyshev expansion on [-kmax, kmax] (Clenshaw in y = 2 (k/kmax)^2 - 1).
__device__ __forceinline__ double ft_eval(const double *c, double k, double kmax)
{
    double x = k / kmax;
    double y = 2.0 * x * x - 1.0;
    double b1 = 0.0, b2 = 0.0;
    for (int n = kNFT - 1; n >= 1; --n) {
        double b0 = fma(2.0 * y, b1, c[n] - b2);
        b2 = b1;
        b1 = b0;
    }
    return fma(y, b1, c[0] - b2);
}

// sin and cos of v: v = n pi/2 + r by Cody-Waite reduction (pi/2 split in
// three parts, the first two with 30 significant bits so that n P1 and n P2
// are exact for |n| < 2^22), Taylor polynomials of sin and cos on
// |r| <= pi/4 (coefficients (-1)^k / j!, truncation below 1e-19), and the
// quadrant of n.  Arguments beyond 2^21 go to the library routine.
__device__ __forceinline__ void sincos_cw(double v, double *s, double *c)
{
    if (!(fabs(v) < 2097152.0)) {
        sincos(v, s, c);
        return;
    }
    const double P1 = 0x1.921fb54p+0, P2 = 0x1.10b46118p-30, P3 = 0x1.313198a2e0370p-61;
    double n = rint(v * 0x1.45f306dc9c883p-1);
    double r = fma(-n, P1, v);
    r = fma(-n, P2, r);
    r = fma(-n, P3, r);
    double r2 = r * r;
    double ps = 0x1.952c77030ad4ap-49;
    ps = fma(ps, r2, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, r2, 0x1.6124613a86d09p-33);
    ps = fma(ps, r2, -0x1.ae64567f544e4p-26);
    ps = fma(ps, r2, 0x1.71de3a556c734p-19);
    ps = fma(ps, r2, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, r2, 0x1.1111111111111p-7);
    ps = fma(ps, r2, -0x1.5555555555555p-3);
    double sr = fma(ps * r2, r, r);
    double pc = -0x1.6827863b97d97p-53;
    pc = fma(pc, r2, 0x1.ae7f3e733b81fp-45);
    pc = fma(pc, r2, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, r2, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, r2, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, r2, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, r2, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, r2, 0x1.5555555555555p-5);
    pc = fma(pc, r2, -0.5);
    double cr = fma(pc, r2, 1.0);
    int q = (int)((long long)n & 3);
    double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// Kernel values at the kw + 1 grid points l = lf + k of a point x, with
// lf = ceil((x - a1) / h): psi_k(f) = ES((2 / kw) (k - kw / 2 + f)) for the
// cell offset f = (lf h - x) / h + kw / 2 in [0, 1), each tap a Chebyshev
// series of degree kEsDeg in y = 2 f - 1 (coefficients per (kw, beta) in the
// kernel table, after the kNFT transform coefficients).
constexpr int kEsDeg = 14;
constexpr int kEsN = kEsDeg + 1;
template <typename R>
__device__ __forceinline__ void cheb_basis(R y, R (&t)[kEsN])
{
    t[0] = (R)1;
    t[1] = y;
    const R y2 = y + y;
#pragma unroll
    for (int n = 2; n < kEsN; ++n) t[n] = fma(y2, t[n - 1], -t[n - 2]);
}
template <typename R>
__device__ __forceinline__ R es_tap(const double *__restrict__ c, const R (&t)[kEsN])
{
    R s = (R)0;
#pragma unroll
    for (int n = kEsN - 1; n >= 0; --n) s = fma((R)c[n], t[n], s);
    return s;
}

// the same tap by Clenshaw's recurrence (no basis array: few registers)
template <typename R>
__device__ __forceinline__ R es_tap_cl(const double *__restrict__ c, R y)
{
    R b1 = (R)0, b2 = (R)0;
    const R y2 = y + y;
#pragma unroll
    for (int n = kEsN - 1; n >= 1; --n) {
        R b0 = fma(y2, b1, (R)c[n] - b2);
        b2 = b1;
        b1 = b0;
    }
    return fma(y, b1, (R)c[0] - b2);
}

__device__ __forceinline__ double es_kernel(double u, double beta)
{
    double t = 1.0 - u * u;
    return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

__device__ __forceinline__ float es_kernel_f(float u, float beta)
{
    float t = 1.0f - u * u;
    return t > 0.0f ? __expf(beta * (__fsqrt_rn(t) - 1.0f)) : 0.0f;
}

// ------------------------------------------------------------------------
__global__ void k_zero(const TfDesc *__restrict__ tfs, const ZeroItem *__restrict__ items)
{
    ZeroItem it = items[blockIdx.x];
    const TfDesc &T = tfs[it.tf];
    double *g = T.gs + (long long)it.col * T.ld_gs;
    for (int q = threadIdx.x; q < it.count; q += blockDim.x) g[it.start + q] = 0.0;
}

__global__ void k_fill0(double *__restrict__ p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// ------------------------------------------------------------------------
// Near-field direct sums, reference add_dir_batch (src/expansions.c:78-97).
// A tile is (rows of one direct box) x (<= 256 sources) x (<= 16 columns);
// sources and their coefficients are staged in LDS, `tpr` lanes share a row
// and are reduced with cross-lane shuffles.
constexpr int kDirTJ = 256;
constexpr int kDirBM = 16;

// J_nu on [0, jtab_n) as piecewise Chebyshev series of degree kJDeg on unit
// intervals; half-integer orders tabulate J_nu(x) / sqrt(x), which is entire.
constexpr int kJDeg = 15;

__device__ __forceinline__ double jtab_eval(const double *__restrict__ tab, double x)
{
    int iv = (int)x;
    double t = 2.0 * (x - (double)iv) - 1.0;
    const double *c = tab + (long long)iv * (kJDeg + 1);
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int n = kJDeg; n >= 1; --n) {
        double b0 = fma(2.0 * t, b1, c[n] - b2);
        b2 = b1;
        b1 = b0;
    }
    return fma(t, b1, c[0] - b2);
}

// Table build: one block of kJDeg + 1 threads per unit interval; values of
// the exact routine at Chebyshev nodes, then the discrete cosine transform.
__global__ void k_jtab_build(int order, int half, int nint, double *__restrict__ tab)
{
    __shared__ double f[kJDeg + 1];
    int iv = blockIdx.x, k = threadIdx.x;
    constexpr int N = kJDeg + 1;
    double tk = cos(kPi * ((double)k + 0.5) / N);
    double x = (double)iv + 0.5 * (tk + 1.0);
    double v = half ? jn_half(order, x, 0.0) : jn_int(order, x, 0.0);
    if (half) v = x > 0.0 ? v / sqrt(x) : 0.0;
    f[k] = v;
    __syncthreads();
    double acc = 0.0;
    for (int q = 0; q < N; ++q) acc += f[q] * cos(kPi * (double)k * ((double)q + 0.5) / N);
    tab[(long long)iv * N + k] = acc * (k == 0 ? 1.0 : 2.0) / N;
}

__global__ __launch_bounds__(256) void k_dir(const TfDesc *__restrict__ tfs,
                                             const DirItem *__restrict__ items,
                                             const int2 *__restrict__ rng,
                                             const double *__restrict__ jtab)
{
    constexpr int LDC = kDirBM + 1;                 // padded row: no bank conflicts
    __shared__ double s_r[kDirTJ];
    __shared__ double s_c[kDirTJ * LDC];
    DirItem it = items[blockIdx.x];
    const TfDesc &T = tfs[it.tf];
    int nb = min(kDirBM, T.B - it.col0);
    int tpr = it.tpr;
    int row = threadIdx.x / tpr, lane = threadIdx.x % tpr;
    bool active = row < it.ni;
    double acc[kDirBM];
#pragma unroll
    for (int t = 0; t < kDirBM; ++t) acc[t] = 0.0;
    const double *tab = jtab + T.jtab_off;
    double xmax = (double)T.jtab_n;
    const double w = active ? T.w[it.i0 + row] : 0.0;
    for (int q = 0; q < it.rcnt; ++q) {
        const int2 rq = rng[it.roff + q];
        for (int c = 0; c < rq.y; c += kDirTJ) {
            const int nj = min(kDirTJ, rq.y - c), j0 = rq.x + c;
            __syncthreads();   // the previous chunk's readers are done
            {
                // one source per thread (nj <= kDirTJ = blockDim), its columns loaded together
                const int j = threadIdx.x;
                const bool ok = j < nj;
                double cv[kDirBM];
#pragma unroll
                for (int t = 0; t < kDirBM; ++t)
                    cv[t] = (ok && t < nb) ? T.cs[(long long)(it.col0 + t) * T.ld_cs + j0 + j] : 0.0;
                s_r[j] = ok ? T.r[j0 + j] : 0.0;
#pragma unroll
                for (int t = 0; t < kDirBM; ++t) s_c[j * LDC + t] = cv[t];
            }
            __syncthreads();
            if (active) {
                for (int j = lane; j < nj; j += tpr) {
                    double x = w * s_r[j];
                    double J;
                    if (x >= 0.0 && x < xmax) {
                        J = jtab_eval(tab, x);
                        if (T.half) J *= sqrt(x);
                    } else {
                        J = besselj_wr(T.order, T.half, w, s_r[j]);
                    }
#pragma unroll
                    for (int t = 0; t < kDirBM; ++t) acc[t] = fma(J, s_c[j * LDC + t], acc[t]);
                }
            }
        }
    }
    for (int off = tpr >> 1; off > 0; off >>= 1) {
#pragma unroll
        for (int t = 0; t < kDirBM; ++t) acc[t] += __shfl_xor(acc[t], off, 64);
    }
    if (active && lane == 0) {
        for (int t = 0; t < nb; ++t)
            atomicAdd(&T.gs[(long long)(it.col0 + t) * T.ld_gs + it.i0 + row], T.sign * acc[t]);
    }
}

// ------------------------------------------------------------------------
// Asymptotic regime, reference add_asy_batch (src/expansions.c:483-559):
//   g[i,t] += sum_e A_e(w_i) Re/Im( e^{i phi} F_e,t(w_i) ),
//   F_e,t(w) = sum_j c_j,t r_j^{-e-1/2} e^{i w r_j},  e = 2l + half,
// i.e. one type-3 NUFFT with E*B complex channels per box (the reference
// issues 2(K+1) separate FINUFFT calls).  The type-3 transform (Lee &
// Greengard 2005) with the exponential-of-semicircle kernel (Barnett,
// Magland & af Klinteberg 2019):
//   x = r - C, s = w - D, v_j = u_j e^{i D x_j}
//   H_l  = sum_j v_j psi(l h - x_j)                    (spread, this kernel)
//   T_m  = sum_l H_l / phihat(l h) e^{2 pi i l m / Nf}  (rocFFT, backward)
//   G(s) = h Delta sum_m phi(s - m Delta) T_m / psihat(s)   (interp kernel)
//   F(w) = e^{i w C} G(s)
// Source weights are normalised by rho = r_j0 (the box's smallest r):
// u_e = c (rho/r)^e r^{-1/2}; the target side carries (rho w)^{-e} w^{-1/2}.

constexpr int kSpreadQ = 128;      // sources staged per LDS pass
constexpr int kTapMax = 17;        // kernel width + 1
constexpr int kInterpWin = 64;     // grid positions staged per target tile

template <typename R> struct Cplx;
template <> struct Cplx<double> { typedef double2 T; };
template <> struct Cplx<float> { typedef float2 T; };

// Blocks go to the 8 XCDs (each with its own L2) round robin in launch
// order.  Item index of this block such that within each group of 8 K blocks
// every XCD takes K consecutive items (neighbouring items read neighbouring
// grid windows); blocks of the last, partial group keep their own index.
template <int K> __device__ __forceinline__ int xcd_item()
{
    const int b = blockIdx.x;
    if (b >= (int)(gridDim.x - gridDim.x % (8 * K))) return b;
    return b - b % (8 * K) + K * (b & 7) + (b >> 3) % K;
}

__device__ __forceinline__ double es_eval(double u, double beta) { return es_kernel(u, beta); }
__device__ __forceinline__ float es_eval(float u, float beta) { return es_kernel_f(u, beta); }

__device__ __forceinline__ void grid_add(double2 *p, double x, double y)
{
    atomicAdd(&p->x, x);
    atomicAdd(&p->y, y);
}
__device__ __forceinline__ void grid_add(float2 *p, float x, float y)
{
    atomicAdd(&p->x, x);
    atomicAdd(&p->y, y);
}

// Zero the (e, t) rows of one position tile that several spread items add into.
template <typename R>
__global__ void k_grid_zero(const TfDesc *__restrict__ tfs, const AsyBoxD *__restrict__ boxes,
                            const AsyItem *__restrict__ items, typename Cplx<R>::T *__restrict__ grid)
{
    AsyItem it = items[blockIdx.x];
    AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    int tb = it.tb;
    int nrow = bx.E * tb;
    for (int q = threadIdx.x; q < nrow * it.count; q += blockDim.x) {
        int row = q / it.count, p = it.start + q % it.count;
        int e = row / tb, t = it.col0 + row % tb;
        if (t >= T.B) continue;
        typename Cplx<R>::T z;
        z.x = (R)0;
        z.y = (R)0;
        grid[bx.goff + (long long)(e * T.B + t) * bx.Nf + p] = z;
    }
}

// R is the arithmetic type of the accumulation (double, or float in mixed
// mode).  Source phases e^{i D x}, kernel offsets l h - x and the
// deconvolution factor are FP64 always.  An item owns a tile of grid
// positions, a column chunk (tb columns) and a source sub-range [jlo, jhi);
// items that split a tile's sources (pad == 1) add atomically into a zeroed
// tile.  Each thread owns one position and two columns; the tb/2 lanes of a
// position read the same tap table entries (LDS broadcast).
template <typename R, int EMAX>
__global__ __launch_bounds__(256) void k_asy_spread(const TfDesc *__restrict__ tfs,
                                                    const AsyBoxD *__restrict__ boxes,
                                                    const AsyItem *__restrict__ items,
                                                    const int2 *__restrict__ ranges,
                                                    typename Cplx<R>::T *__restrict__ grid,
                                                    const double *__restrict__ ft)
{
    typedef typename Cplx<R>::T C2;
    __shared__ C2 s_z[kSpreadQ][kTapMax];        // psi * e^{iDx} r^{-1/2}
    __shared__ C2 s_c[kSpreadQ][8];              // column pairs
    __shared__ double s_x[kSpreadQ], s_br[kSpreadQ], s_bi[kSpreadQ];
    __shared__ R s_ri[kSpreadQ];
    __shared__ int s_pf[kSpreadQ];
    AsyItem it = items[blockIdx.x];
    int2 jr = ranges[blockIdx.x];
    AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    int tb = it.tb;
    int tl2 = tb > 1 ? tb >> 1 : 1;          // lanes per position
    int tl = threadIdx.x % tl2, pl = threadIdx.x / tl2;
    int p = it.start + pl;
    int l = p <= bx.L ? p : p - bx.Nf;
    bool valid = (pl < it.count) && (l >= -bx.L) && (l <= bx.L);
    int t0 = it.col0 + 2 * tl;
    bool active = valid && (t0 < T.B);
    int kw = T.kw;
    double inv_a1 = 1.0 / bx.a1;
    R a0r[EMAX], a0i[EMAX], a1r[EMAX], a1i[EMAX];
#pragma unroll
    for (int e = 0; e < EMAX; ++e) { a0r[e] = a0i[e] = a1r[e] = a1i[e] = (R)0; }
    int E = bx.E;
    for (int c0 = jr.x; c0 < jr.y; c0 += kSpreadQ) {
        int cn = min(kSpreadQ, jr.y - c0);
        for (int q = threadIdx.x; q < cn; q += 256) {
            double rj = T.r[c0 + q];
            double x = rj - bx.C;
            double sn, cs;
            sincos_cw(bx.D * x, &sn, &cs);
            double rs = 1.0 / sqrt(rj);
            int lf = (int)ceil((x - bx.a1) / bx.h);
            s_x[q] = x;
            s_br[q] = cs * rs;
            s_bi[q] = sn * rs;
            s_pf[q] = lf;
            s_ri[q] = (R)(bx.rho / rj);
        }
        for (int q = threadIdx.x; q < cn * 8; q += 256) {
            int jq = q >> 3, pr = q & 7;
            int ta = it.col0 + 2 * pr, tb2 = ta + 1;
            C2 v;
            v.x = ta < T.B && 2 * pr < tb ? (R)T.cs[(long long)ta * T.ld_cs + c0 + jq] : (R)0;
            v.y = tb2 < T.B && 2 * pr + 1 < tb ? (R)T.cs[(long long)tb2 * T.ld_cs + c0 + jq] : (R)0;
            s_c[jq][pr] = v;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < cn * kTapMax; q += 256) {
            int jq = q / kTapMax, k = q % kTapMax;
            double psi = 0.0;
            if (k <= kw) {
                double u = ((double)(s_pf[jq] + k) * bx.h - s_x[jq]) * inv_a1;
                psi = (double)es_eval((R)u, (R)T.beta);
            }
            C2 z;
            z.x = (R)(psi * s_br[jq]);
            z.y = (R)(psi * s_bi[jq]);
            s_z[jq][k] = z;
        }
        __syncthreads();
        if (active) {
            // sources of this chunk whose taps reach l: lf in [l - kw, l] (lf ascending)
            int a = 0, b = cn;
            while (a < b) { int mid = (a + b) >> 1; if (s_pf[mid] >= l - kw) b = mid; else a = mid + 1; }
            int qa = a;
            b = cn;
            while (a < b) { int mid = (a + b) >> 1; if (s_pf[mid] > l) b = mid; else a = mid + 1; }
            int qb = a;
            for (int q = qa; q < qb; ++q) {
                C2 z = s_z[q][l - s_pf[q]];
                C2 cc = s_c[q][tl];
                R ri = s_ri[q];
                R z0r = z.x * cc.x, z0i = z.y * cc.x, z1r = z.x * cc.y, z1i = z.y * cc.y;
                R rp = (R)1;
#pragma unroll
                for (int e = 0; e < EMAX; ++e) {
                    if (e < E) {
                        a0r[e] = fma(z0r, rp, a0r[e]);
                        a0i[e] = fma(z0i, rp, a0i[e]);
                        a1r[e] = fma(z1r, rp, a1r[e]);
                        a1i[e] = fma(z1i, rp, a1i[e]);
                        rp *= ri;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (pl < it.count && t0 < T.B) {
        double dec = 0.0;
        if (valid) dec = 1.0 / (bx.a2 * ft_eval(ft + T.ft_off, (double)l * bx.h * bx.a2, T.kmax));
        bool two = (t0 + 1 < T.B) && (2 * tl + 1 < tb);
        for (int e = 0; e < E; ++e) {
            C2 *d0 = grid + bx.goff + ((long long)(e * T.B + t0)) * bx.Nf + p;
            R v0r = (R)((double)a0r[e] * dec), v0i = (R)((double)a0i[e] * dec);
            R v1r = (R)((double)a1r[e] * dec), v1i = (R)((double)a1i[e] * dec);
            if (it.pad) {
                if (valid) {
                    grid_add(d0, v0r, v0i);
                    if (two) grid_add(d0 + bx.Nf, v1r, v1i);
                }
            } else {
                C2 v;
                v.x = v0r;
                v.y = v0i;
                *d0 = v;
                if (two) {
                    v.x = v1r;
                    v.y = v1i;
                    d0[bx.Nf] = v;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// MFMA formulations of the asymptotic stage for dense boxes.
//
// 16x16x4 matrix cores, one wave (64 lanes) per item.  Operand layouts of
// both builtins (checked on gfx950): A[i][k] in lane i + 16 k, B[k][j] in
// lane j + 16 k; the 16x16 result sits in 4 registers per lane, row
// l/16 + 4 v (f64) or 4 (l/16) + v (f32), column l % 16.
typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

template <typename R> struct Mfma;
template <> struct Mfma<double> {
    typedef mf_d4 V;
    static __device__ __forceinline__ V mma(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
    typedef mf_f4 V;
    static __device__ __forceinline__ V mma(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// ------------------------------------------------------------------------
// Local expansion, projection half: coef[l][t] += sum_j B_l(x_j) c_j,t with
// x_j = r_j / r_end.  Integer orders follow the reference's Wimp expansion
// (src/expansions.c:242-273): even nu -> eps_l T_2l(x), odd nu -> 2 T_2l+1(x).
// Half-integer orders (Chebyshev interpolation of j_n, this build) use
// T_2l+p(x), p = n mod 2, with weights c_j sqrt(r_j).
//
// The contraction is a (K+1) x count x 16 matrix product per item, run on
// the FP64 matrix cores: A[l][q] = B_l(x_q) (NLT row tiles of 16), B[q][t] =
// c_q,t, accumulated in registers over every source of the item; a pass
// stages kLocSub sources (basis by the three-term recurrence, one source per
// lane; coefficients loaded column by column, coalesced) in LDS with odd row
// strides, the 4 waves take interleaved K steps of 4 sources, and the 4
// partial tiles are summed through LDS before one atomic add per output.
constexpr int kLocBM = 16;     // columns per item
constexpr int kLocSub = 64;    // sources staged per pass
constexpr int kLocMaxNlt = 6;  // row tiles: K_loc <= 95

template <int NLT> struct LocProjLds {
    static constexpr int LT = 16 * NLT + 1;
    static constexpr size_t stage = (size_t)kLocSub * (LT + kLocBM + 1);
    static constexpr size_t red = (size_t)4 * 16 * NLT * (kLocBM + 1);
    static constexpr size_t bytes = (stage > red ? stage : red) * sizeof(double);
};

template <int NLT>
__global__ __launch_bounds__(256) void k_loc_proj(const TfDesc *__restrict__ tfs,
                                                  const LocBoxD *__restrict__ boxes,
                                                  const LocItem *__restrict__ items,
                                                  double *__restrict__ coef)
{
    constexpr int LT = LocProjLds<NLT>::LT, LC = kLocBM + 1;
    extern __shared__ double sm[];
    double *sT = sm;                          // [kLocSub][LT]
    double *sC = sm + (size_t)kLocSub * LT;   // [kLocSub][LC]
    LocItem it = items[blockIdx.x];
    LocBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    const int K = T.K_loc, par = T.order & 1;
    const int nb = min(kLocBM, T.B - it.col0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    mf_d4 d[NLT];
#pragma unroll
    for (int lt = 0; lt < NLT; ++lt)
        for (int v = 0; v < 4; ++v) d[lt][v] = 0.0;
    for (int s0 = 0; s0 < it.count; s0 += kLocSub) {
        const int n = min(kLocSub, it.count - s0);
        __syncthreads();
        if (tid < kLocSub) {
            double *row = sT + (size_t)tid * LT;
            if (tid < n) {
                double x = T.r[bx.j0 + it.start + s0 + tid] / bx.r_end;
                x = fmin(1.0, fmax(-1.0, x));
                double tm1 = 1.0, t0 = x;   // T_0, T_1
                row[0] = par ? 2.0 * x : 1.0;
                int deg = 1;
                for (int l = 1; l < 16 * NLT; ++l) {
                    if (l <= K) {
                        for (int target = 2 * l + par; deg < target; ++deg) {
                            double tn = 2.0 * x * t0 - tm1;
                            tm1 = t0;
                            t0 = tn;
                        }
                        row[l] = 2.0 * t0;
                    } else {
                        row[l] = 0.0;
                    }
                }
            } else {
                for (int l = 0; l < 16 * NLT; ++l) row[l] = 0.0;
            }
        }
        // coefficients: lanes run over sources, 256 / kLocSub columns at a time
        for (int e = tid; e < kLocSub * kLocBM; e += 256) {
            int q = e % kLocSub, t = e / kLocSub;
            double c = 0.0;
            if (q < n && t < nb) {
                int j = bx.j0 + it.start + s0 + q;
                c = T.cs[(long long)(it.col0 + t) * T.ld_cs + j];
                if (T.half) c *= sqrt(T.r[j]);
            }
            sC[q * LC + t] = c;
        }
        __syncthreads();
        const int ia = lane & 15, ka = lane >> 4;
        for (int st = wave; 4 * st < n; st += 4) {
            const int q = 4 * st + ka;
            const double bv = sC[q * LC + ia];
#pragma unroll
            for (int lt = 0; lt < NLT; ++lt)
                d[lt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sT[q * LT + 16 * lt + ia], bv, d[lt],
                                                            0, 0, 0);
        }
    }
    // sum the 4 waves' tiles: red[wave][16 NLT][LC]
    __syncthreads();
    double *red = sm;
#pragma unroll
    for (int lt = 0; lt < NLT; ++lt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int row = 16 * lt + (lane >> 4) + 4 * v;
            red[((size_t)wave * 16 * NLT + row) * LC + (lane & 15)] = d[lt][v];
        }
    __syncthreads();
    for (int e = tid; e < 16 * NLT * 16; e += 256) {
        int row = e >> 4, t = e & 15;
        if (row > K || t >= nb) continue;
        double acc = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) acc += red[((size_t)w * 16 * NLT + row) * LC + t];
        atomicAdd(&coef[bx.coef_off + (long long)row * T.B + it.col0 + t], acc);
    }
}

// Local expansion, evaluation half (reference src/expansions.c:275-336):
// integer orders compute J_0..J_{l0+K+odd}(z), z = w r_end / 2, by one
// Miller sweep and form b_l = J_{l0+odd+l}(z) J_{|l0-l|}(z) (sign (-1) for
// l > l0 with l - l0 odd); half-integer orders sample j_n(w r_end x_q) at
// the K+1 positive Chebyshev nodes and map them to coefficients with the
// precomputed matrix cheb[].  Then g[i,t] += sum_l b_l coef[l][t] for every
// column t.  One thread per target.  k_loc_eval1: one local box per item;
// the sums go out four columns at a time.  k_loc_eval: every box that
// covers the item's targets (about 2.6 at the clustered workload), looped
// over with up to BMAX column sums in registers, so that each output element
// gets one atomic add per item instead of one per box (the registers cost
// occupancy, so transforms whose targets mostly meet one box keep the first
// form).  The per-target scratch lives in LDS as
// [slot][thread] (slots = loc_slots()), so lanes touch consecutive words;
// a box's coefficients are the same for the whole block and are read as
// uniform loads.
__host__ __device__ inline int loc_slots(int order, int half, int K)
{
    return half ? 2 * (K + 1) : (order >> 1) + K + (order & 1) + 1;
}

__global__ __launch_bounds__(256) void k_loc_eval1(const TfDesc *__restrict__ tfs, const LocBoxD *__restrict__ boxes,
                           const LocItem *__restrict__ items, const double *__restrict__ coef,
                           const double *__restrict__ cheb)
{
    extern __shared__ double sm[];
    LocItem it = items[blockIdx.x];
    LocBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    const int K = T.K_loc;
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (tid >= it.count) return;
    double *S = sm + tid;   // slot k at S[k * nthr]
    const int i = bx.i0 + it.start + tid;
    const double w = T.w[i];
    double scale = 1.0;
    int boff;               // slot of b_0 (b_l at slot boff + l)
    if (!T.half) {
        const int l0 = T.order >> 1, odd = T.order & 1;
        const int maxo = l0 + K + odd;
        const double z = w * bx.r_end * 0.5;
        if (z == 0.0) {
            for (int k = 0; k <= maxo; ++k) S[k * nthr] = (k == 0) ? 1.0 : 0.0;
        } else {
            double m = (double)(maxo > (int)z ? maxo : (int)z);
            int N = (int)(m + 20.0 + 4.0 * sqrt(m + 1.0) + 6.0 * cbrt(m + 1.0));
            N += N & 1;
            double jp1 = 0.0, j = 1e-280, sum = 0.0, tx = 2.0 / z;
            for (int k = N; k >= 1; --k) {
                double jm1 = (double)k * tx * j - jp1;
                jp1 = j;
                j = jm1;
                int km1 = k - 1;
                if (km1 <= maxo) S[km1 * nthr] = j;
                if (km1 > 0 && !(km1 & 1)) sum += 2.0 * j;
                if (fabs(j) > 1e200) {
                    j *= 1e-200; jp1 *= 1e-200; sum *= 1e-200;
                    for (int q = km1; q <= maxo; ++q) S[q * nthr] *= 1e-200;
                }
            }
            double inv = 1.0 / (sum + j);
            for (int k = 0; k <= maxo; ++k) S[k * nthr] *= inv;
        }
        // b_l replaces J_{l0+odd+l}; taken from the top down, every J still
        // needed sits below the slots already replaced
        for (int l = K; l >= 0; --l) {
            int lo = l0 - l;
            lo = lo < 0 ? -lo : lo;
            double v = S[(l0 + odd + l) * nthr] * S[lo * nthr];
            if (l > l0 && ((l - l0) & 1)) v = -v;
            S[(l0 + odd + l) * nthr] = v;
        }
        boff = l0 + odd;
    } else {
        // f_q = j_n(z x_q), x_q = cos(pi (q + 1/2) / (2K + 2)), q = 0..K, at
        // slots K+1+q; b_l at slots l
        const double z = w * bx.r_end;
        for (int q = 0; q <= K; ++q) {
            double xq = cos(kPi * ((double)q + 0.5) / (double)(2 * K + 2));
            double hi = z * xq, lo = fma(z, xq, -hi);
            S[(K + 1 + q) * nthr] = sph_jn(T.order, hi, lo);
        }
        const double *M = cheb + T.cheb_off;
        for (int l = 0; l <= K; ++l) {
            double a = 0.0;
            for (int q = 0; q <= K; ++q) a = fma(M[l * (K + 1) + q], S[(K + 1 + q) * nthr], a);
            S[l * nthr] = a;
        }
        boff = 0;
        scale = sqrt(2.0 * w / kPi);
    }
    const double sc = T.sign * scale;
    const double *cf = coef + bx.coef_off;
    const int B = T.B;
    for (int t0 = 0; t0 < B; t0 += 4) {
        const int nt = min(4, B - t0);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (int l = 0; l <= K; ++l) {
            double bl = S[(boff + l) * nthr];
            const double *cr = cf + (long long)l * B + t0;
            a0 = fma(bl, cr[0], a0);
            if (nt > 1) a1 = fma(bl, cr[1], a1);
            if (nt > 2) a2 = fma(bl, cr[2], a2);
            if (nt > 3) a3 = fma(bl, cr[3], a3);
        }
        double *g = T.gs + (long long)t0 * T.ld_gs + i;
        atomicAdd(g, sc * a0);
        if (nt > 1) atomicAdd(g + T.ld_gs, sc * a1);
        if (nt > 2) atomicAdd(g + 2 * (long long)T.ld_gs, sc * a2);
        if (nt > 3) atomicAdd(g + 3 * (long long)T.ld_gs, sc * a3);
    }
}

template <int BMAX>
__global__ __launch_bounds__(256) void k_loc_eval(const TfDesc *__restrict__ tfs, const LocBoxD *__restrict__ boxes,
                           const LocEvalItem *__restrict__ items,
                           const int *__restrict__ boxlist, const double *__restrict__ coef,
                           const double *__restrict__ cheb)
{
    extern __shared__ double sm[];
    const LocEvalItem it = items[blockIdx.x];
    const TfDesc &T = tfs[it.tf];
    const int K = T.K_loc;
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (tid >= it.count) return;
    double *S = sm + tid;   // slot k at S[k * nthr]
    const int i = it.i0 + tid;
    const double w = T.w[i];
    const int B = T.B, nb = min(BMAX, B - it.col0);
    double acc[BMAX];
#pragma unroll
    for (int t = 0; t < BMAX; ++t) acc[t] = 0.0;
    for (int q = 0; q < it.rcnt; ++q) {
        const LocBoxD bx = boxes[boxlist[it.roff + q]];
        int boff;               // slot of b_0 (b_l at slot boff + l)
        if (!T.half) {
            const int l0 = T.order >> 1, odd = T.order & 1;
            const int maxo = l0 + K + odd;
            const double z = w * bx.r_end * 0.5;
            if (z == 0.0) {
                for (int k = 0; k <= maxo; ++k) S[k * nthr] = (k == 0) ? 1.0 : 0.0;
            } else {
                double m = (double)(maxo > (int)z ? maxo : (int)z);
                int N = (int)(m + 20.0 + 4.0 * sqrt(m + 1.0) + 6.0 * cbrt(m + 1.0));
                N += N & 1;
                double jp1 = 0.0, j = 1e-280, sum = 0.0, tx = 2.0 / z;
                for (int k = N; k >= 1; --k) {
                    double jm1 = (double)k * tx * j - jp1;
                    jp1 = j;
                    j = jm1;
                    int km1 = k - 1;
                    if (km1 <= maxo) S[km1 * nthr] = j;
                    if (km1 > 0 && !(km1 & 1)) sum += 2.0 * j;
                    if (fabs(j) > 1e200) {
                        j *= 1e-200; jp1 *= 1e-200; sum *= 1e-200;
                        for (int qq = km1; qq <= maxo; ++qq) S[qq * nthr] *= 1e-200;
                    }
                }
                double inv = 1.0 / (sum + j);
                for (int k = 0; k <= maxo; ++k) S[k * nthr] *= inv;
            }
            // b_l replaces J_{l0+odd+l}; taken from the top down, every J still
            // needed sits below the slots already replaced
            for (int l = K; l >= 0; --l) {
                int lo = l0 - l;
                lo = lo < 0 ? -lo : lo;
                double v = S[(l0 + odd + l) * nthr] * S[lo * nthr];
                if (l > l0 && ((l - l0) & 1)) v = -v;
                S[(l0 + odd + l) * nthr] = v;
            }
            boff = l0 + odd;
        } else {
            // f_q = j_n(z x_q), x_q = cos(pi (q + 1/2) / (2K + 2)), q = 0..K, at
            // slots K+1+q; b_l at slots l
            const double z = w * bx.r_end;
            for (int qq = 0; qq <= K; ++qq) {
                double xq = cos(kPi * ((double)qq + 0.5) / (double)(2 * K + 2));
                double hi = z * xq, lo = fma(z, xq, -hi);
                S[(K + 1 + qq) * nthr] = sph_jn(T.order, hi, lo);
            }
            const double *M = cheb + T.cheb_off;
            for (int l = 0; l <= K; ++l) {
                double a = 0.0;
                for (int qq = 0; qq <= K; ++qq) a = fma(M[l * (K + 1) + qq], S[(K + 1 + qq) * nthr], a);
                S[l * nthr] = a;
            }
            boff = 0;
        }
        const double *cf = coef + bx.coef_off + it.col0;
        for (int l = 0; l <= K; ++l) {
            const double bl = S[(boff + l) * nthr];
            const double *cr = cf + (long long)l * B;
#pragma unroll
            for (int t = 0; t < BMAX; ++t)
                if (t < nb) acc[t] = fma(bl, cr[t], acc[t]);
        }
    }
    const double sc = T.sign * (T.half ? sqrt(2.0 * w / kPi) : 1.0);
    double *g = T.gs + (long long)it.col0 * T.ld_gs + i;
#pragma unroll
    for (int t = 0; t < BMAX; ++t)
        if (t < nb) atomicAdd(g + (long long)t * T.ld_gs, sc * acc[t]);
}

// Spread, source-dense boxes, on the matrix cores.  A block takes a range
// of consecutive sources of a box (a piece) and columns [col0, col0 + tb);
// with nsub = 16 / tb:
//   G[p][e][t] = sum_q psi(l_p h - x_q) u_q^(e) c[q][t],  u^(e) = e^{iDx} r^{-1/2} (rho/r)^e
// as 16 x 16 matrix products over q, one per (exponent group eg < ceil(E /
// nsub), real or imaginary part): rows j = (t, e2) with e2 < nsub and e =
// eg nsub + e2, A[j][q] = c[q][t] (rho/r_q)^{e2}; columns = 16 grid
// positions (a window that slides with the sources), B[q][p] = psi(l_p h -
// x_q) Re or Im u_q^(eg nsub).  Each source is multiplied once, with the
// kw + 1 of the 16 columns its kernel reaches.  The block's consumer waves
// split the products between them and each runs every source, so that all
// of them stay busy on every staged chunk.  Per source, the phase, powers,
// kernel values and the A operand are computed once per block and shared
// through LDS.
constexpr int kSQ = 64;           // sources staged per LDS pass
constexpr int kSP = 16;           // accumulator columns: a window of grid positions

// Sources [piece.x, piece.y) of a box; the item writes grid positions l in
// [wr.x, wr.y], with plain stores in [ex.x, ex.y] (no other item reaches
// them) and atomic adds elsewhere (zeroed beforehand).
struct SpreadRange { int2 piece, wr, ex, pad; };

// Row strides are padded so that both the staging writes (one source per
// lane) and the product reads (4 sources x 4 positions or exponents per
// instruction) spread over the LDS banks.
template <typename R, int EMAX> struct SpreadLds {
    typedef typename Cplx<R>::T C2;
    R psi[kSQ][kSP + 5];          // kernel value of source q at position l: column l mod 16
    C2 u[kSQ][EMAX + 1];          // e^{iDx} r^{-1/2} (rho/r)^e
    R bop[kSQ][17];               // A operand: c[q][t] (rho/r)^{e2}, row j = (t, e2)
    int lf[kSQ];                  // first kernel tap of source q
};

// Products of one staged chunk: KC of the block's products, accumulated over
// the chunk's sources 4 per step; uoff[k] indexes product k's part of u in a
// source's staged row.  Rows past the chunk's end are zero, so the loop needs
// no masks and the accumulators stay in registers.
template <typename R, int EMAX, int KC, int KA>
__device__ __forceinline__ void spread_chunk(const SpreadLds<R, EMAX> &L, int qhi, int lane,
                                             const int (&uoff)[KA],
                                             typename Mfma<R>::V (&acc)[KA])
{
    const int ka = lane >> 4, ja = lane & 15;
    int s = 0;
    // two steps per iteration: the second step's operands load while the
    // first step's products run
    for (; s + 8 <= qhi; s += 8) {
        const int q0 = s + ka, q1 = q0 + 4;
        const R a0 = L.bop[q0][ja], a1 = L.bop[q1][ja];
        const R p0 = L.psi[q0][ja], p1 = L.psi[q1][ja];
        const R *u0 = (const R *)&L.u[q0][0], *u1 = (const R *)&L.u[q1][0];
        R b0[KC], b1[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            b0[k] = p0 * u0[uoff[k]];
            b1[k] = p1 * u1[uoff[k]];
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) acc[k] = Mfma<R>::mma(a0, b0[k], acc[k]);
#pragma unroll
        for (int k = 0; k < KC; ++k) acc[k] = Mfma<R>::mma(a1, b1[k], acc[k]);
    }
    for (; s < qhi; s += 4) {
        const int q = s + ka;
        const R a = L.bop[q][ja];
        const R ps = L.psi[q][ja];
        const R *uq = (const R *)&L.u[q][0];
#pragma unroll
        for (int k = 0; k < KC; ++k) acc[k] = Mfma<R>::mma(a, ps * uq[uoff[k]], acc[k]);
    }
}

// One step (sources s .. s + 3) restricted to the sources s + lo .. s + hi - 1.
template <typename R, int EMAX, int KC, int KA>
__device__ __forceinline__ void spread_step(const SpreadLds<R, EMAX> &L, int s, int lo, int hi,
                                            int lane, const int (&uoff)[KA],
                                            typename Mfma<R>::V (&acc)[KA])
{
    const int ka = lane >> 4, ja = lane & 15;
    const int q = s + ka;
    const R a = L.bop[q][ja];
    const R ps = (ka >= lo && ka < hi) ? L.psi[q][ja] : (R)0;
    const R *uq = (const R *)&L.u[q][0];
#pragma unroll
    for (int k = 0; k < KC; ++k) acc[k] = Mfma<R>::mma(a, ps * uq[uoff[k]], acc[k]);
}

// Global loads of one producer thread for one chunk (its source's r; for
// waves 2-3 also 8 coefficients), issued a chunk ahead of their use.
template <bool CV> struct SpreadIn { double r; double cv[8]; };
template <> struct SpreadIn<false> { double r; };

template <bool CV>
__device__ __forceinline__ void spread_load(const TfDesc &T, const AsyItem &it, int c0, int cn,
                                            int ptid, SpreadIn<CV> &in)
{
    const int q = ptid & 63;
    in.r = T.r[c0 + min(q, cn - 1)];
    if constexpr (CV) {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int j = (ptid - 128 + 128 * m) >> 6, tc = it.col0 + j % it.tb;
            in.cv[m] = (q < cn && tc < T.B) ? T.cs[(long long)tc * T.ld_cs + c0 + q] : 0.0;
        }
    }
}

// Staging of one chunk of up to kSQ sources by the 4 producer waves
// (pw = producer wave, ptid = its thread index in 0..255) from their loads:
// wave 0 the phases and powers, wave 1 the kernel values at the positions
// the sources reach, waves 2-3 the A operand.  Rows past the chunk's end
// are zero.
template <typename R, int EMAX, bool CV>
__device__ __forceinline__ void spread_stage(const TfDesc &T, const AsyBoxD &bx, const AsyItem &it,
                                             int c0, int cn, int pw, int ptid,
                                             const SpreadIn<CV> &in, const double *__restrict__ ft,
                                             SpreadLds<R, EMAX> &L)
{
    typedef typename Cplx<R>::T C2;
    const int lane = ptid & 63, tb = it.tb, E = bx.E, kw = T.kw;
    if (pw == 0) {
        const int q = lane;
        double rj = q < cn ? in.r : bx.C;
        double x = rj - bx.C;
        double sn, cs;
        sincos_cw(bx.D * x, &sn, &cs);
        double rs = q < cn ? 1.0 / sqrt(rj) : 0.0;
        double br = cs * rs, bi = sn * rs;
        double ri = q < cn ? bx.rho / rj : 0.0, rp = 1.0;
        for (int e = 0; e < E; ++e) {
            C2 u;
            u.x = (R)(br * rp);
            u.y = (R)(bi * rp);
            L.u[q][e] = u;
            rp *= ri;
        }
    } else if (pw == 1) {
        // taps l = lf .. lf + kw at columns l mod 16 (distinct for kw < 16);
        // rows past the chunk's end: zero values, the last source's lf
        const int q = lane;
        const double inv_h = 1.0 / bx.h;
#pragma unroll
        for (int pi = 0; pi < kSP; ++pi) L.psi[q][pi] = (R)0;
        const double x = in.r - bx.C;
        const int lf = (int)ceil((x - bx.a1) * inv_h);
        L.lf[q] = lf;
        if (q < cn) {
            double f = ((double)lf * bx.h - x) * inv_h + 0.5 * kw;
            R tb_[kEsN];
            cheb_basis((R)(2.0 * f - 1.0), tb_);
            const double *tc = ft + T.ft_off + kNFT;
            for (int k = 0; k <= kw; ++k) {
                int l = lf + k;
                if (l >= -bx.L && l <= bx.L) L.psi[q][l & 15] = es_tap(tc + k * kEsN, tb_);
            }
        }
    } else if constexpr (CV) {
        // A operand: 8 (source, row) entries per thread, one source
        const int q = lane;
        const double ri = q < cn ? bx.rho / in.r : 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int j = (ptid - 128 + 128 * m) >> 6, e2 = j / tb;
            double c = in.cv[m];
            if (e2 > 0) {
                double rp = ri;
                for (int e = 1; e < e2; ++e) rp *= ri;
                c *= rp;
            }
            L.bop[q][j] = (R)c;
        }
    }
}

// The producer side of the split schedule: stage chunk 0, then chunk k + 1
// while the consumers multiply chunk k, one block barrier per chunk (as many
// as the consumers pass).  The loads of chunk k + 1 are in flight while
// chunk k is staged and the barrier waits.
template <typename R, int EMAX, bool CV>
__device__ __forceinline__ void spread_produce(const TfDesc &T, const AsyBoxD &bx,
                                               const AsyItem &it, int2 piece, int nchunk, int pw,
                                               int ptid, const double *__restrict__ ft,
                                               SpreadLds<R, EMAX> *L)
{
    SpreadIn<CV> in;
    if (nchunk > 0) {
        spread_load(T, it, piece.x, min(kSQ, piece.y - piece.x), ptid, in);
        spread_stage<R, EMAX>(T, bx, it, piece.x, min(kSQ, piece.y - piece.x), pw, ptid, in, ft,
                              L[0]);
        if (nchunk > 1) {
            const int c1 = piece.x + kSQ;
            spread_load(T, it, c1, min(kSQ, piece.y - c1), ptid, in);
        }
    }
    __syncthreads();
    for (int k = 0; k < nchunk; ++k) {
        if (k + 1 < nchunk) {
            SpreadIn<CV> nx = in;
            const int c1 = piece.x + (k + 1) * kSQ, c2 = c1 + kSQ;
            if (k + 2 < nchunk) spread_load(T, it, c2, min(kSQ, piece.y - c2), ptid, nx);
            spread_stage<R, EMAX>(T, bx, it, c1, min(kSQ, piece.y - c1), pw, ptid, in, ft,
                                  L[(k + 1) & 1]);
            in = nx;
        }
        __syncthreads();
    }
}

// Two schedules of the same work.  SPLIT: a block of 8 waves,
// waves 0-3 (consumers) hold the accumulators, waves 4-7 (producers) stage
// source chunks; chunks are double-buffered in LDS (the consumers multiply
// chunk k out of buffer k mod 2 while the producers stage chunk k + 1 into
// the other; one block barrier per chunk), and the staging code's registers
// never coexist with the accumulators.  Otherwise: a block of 4 waves that
// stage a chunk together, one task each, then multiply it.  KW = products
// per consumer wave (wave w takes products [m0, m0 + mc), mc = KW or KW - 1).
//
// The 16 accumulator columns are a window sliding over the grid: column c
// holds position P(c) = W + ((c - W) mod 16), the one of [W, W + 16) that is
// c mod 16.  Sources come in ascending order (so do their first taps lf);
// before a source's products the window start W moves up to lf + kw - 15 (if
// it is not there already), which keeps W <= lf: the positions it passes
// receive nothing more, and are written out and their columns cleared.
template <typename R, int EMAX, int KW, bool SPLIT>
__device__ __forceinline__ void spread_mfma_body(const TfDesc &T, const AsyBoxD &bx,
                                                 const AsyItem &it, const SpreadRange &rg,
                                                 int nprod, typename Cplx<R>::T *__restrict__ grid,
                                                 const double *__restrict__ ft,
                                                 SpreadLds<R, EMAX> *L)
{
    typedef typename Mfma<R>::V V;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int2 piece = rg.piece;
    const int nchunk = (piece.y - piece.x + kSQ - 1) / kSQ;
    if constexpr (SPLIT) {
        if (wv >= 4) {
            // waves 6-7 hold the coefficients they load, waves 4-5 only r
            if (wv >= 6) spread_produce<R, EMAX, true>(T, bx, it, piece, nchunk, wv - 4, tid - 256, ft, L);
            else spread_produce<R, EMAX, false>(T, bx, it, piece, nchunk, wv - 4, tid - 256, ft, L);
            return;
        }
        __syncthreads();
    }
    const int tb = it.tb, nsub = 16 / tb;
    const int E = bx.E, kw = T.kw;
    // this wave's products: the nprod products split as evenly as possible
    const int base = nprod >> 2, rem = nprod & 3;
    const int m0 = wv * base + min(wv, rem), mc = base + (wv < rem ? 1 : 0);
    int uoff[KW];
    V acc[KW];
#pragma unroll
    for (int k = 0; k < KW; ++k) {
        const int m = min(m0 + k, nprod - 1);
        uoff[k] = 2 * (m >> 1) * nsub + (m & 1);
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[k][v] = (R)0;
    }
    const int col = lane & 15;
    // this lane's result rows j = (tt, e2), packed e2 << 8 | tt (-1: no column)
    int rowpk[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int j = Mfma<R>::row(lane, v), tt = j % tb;
        rowpk[v] = it.col0 + tt < T.B ? ((j / tb) << 8) | tt : -1;
    }
    int W = 0;
    // positions [W, Wn) are final: write this lane's column if it holds one
    auto flush = [&](int Wn) {
        const int P = W + ((col - W) & 15);
        if (P >= Wn) return;
        if (P >= rg.wr.x && P <= rg.wr.y) {
            // (the opaque copies keep the row and address arithmetic here,
            // where it runs once per window move, instead of held in
            // registers across the whole loop)
            int pk[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                pk[v] = rowpk[v];
                asm volatile("" : "+v"(pk[v]));
            }
            const bool excl = P >= rg.ex.x && P <= rg.ex.y;
            const int pos = P >= 0 ? P : P + bx.Nf;
            const double dec = 1.0 / (bx.a2 * ft_eval(ft + T.ft_off, (double)P * bx.h * bx.a2, T.kmax));
#pragma unroll
            for (int k = 0; k < KW; ++k) {
                if (k >= mc) break;
                const int m = m0 + k, eg = m >> 1, part = m & 1;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int e = eg * nsub + (pk[v] >> 8);
                    if (pk[v] < 0 || e >= E) continue;
                    const int t = it.col0 + (pk[v] & 255);
                    R *d = (R *)(grid + bx.goff + ((long long)(e * T.B + t)) * bx.Nf + pos) + part;
                    const R val = (R)((double)acc[k][v] * dec);
                    if (excl) *d = val;
                    else atomicAdd(d, val);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KW; ++k)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[k][v] = (R)0;
    };
    auto advance = [&](int Wn) {
        if (Wn > W) {
            flush(Wn);
            W = Wn;
        }
    };
    for (int k = 0; k < nchunk; ++k) {
        const int c0 = piece.x + k * kSQ, cn = min(kSQ, piece.y - c0);
        if constexpr (!SPLIT) {
            if (wv >= 2) {
                SpreadIn<true> in;
                spread_load(T, it, c0, cn, tid, in);
                spread_stage<R, EMAX>(T, bx, it, c0, cn, wv, tid, in, ft, L[0]);
            } else {
                SpreadIn<false> in;
                spread_load(T, it, c0, cn, tid, in);
                spread_stage<R, EMAX>(T, bx, it, c0, cn, wv, tid, in, ft, L[0]);
            }
            __syncthreads();
        }
        const SpreadLds<R, EMAX> &Lk = L[SPLIT ? (k & 1) : 0];
        const int l0 = __builtin_amdgcn_readfirstlane(Lk.lf[0]);
        const int l1 = __builtin_amdgcn_readfirstlane(Lk.lf[cn - 1]);
        if (k == 0) W = l0 + kw - 15;
        if (l1 - l0 <= 15 - kw) {
            // the whole chunk fits one window
            advance(l1 + kw - 15);
            if (mc == KW) spread_chunk<R, EMAX, KW, KW>(Lk, cn, lane, uoff, acc);
            else if constexpr (KW > 1) {
                if (mc == KW - 1) spread_chunk<R, EMAX, KW - 1, KW>(Lk, cn, lane, uoff, acc);
            }
        } else {
            // step by step; a step whose sources do not fit one window runs
            // in parts, each after the window has moved for its first source
            for (int s = 0; s < cn; s += 4) {
                int lo = 0;
                while (lo < 4) {
                    const int llo = __builtin_amdgcn_readfirstlane(Lk.lf[min(s + lo, cn - 1)]);
                    advance(llo + kw - 15);
                    int hi = lo + 1;
                    while (hi < 4 &&
                           __builtin_amdgcn_readfirstlane(Lk.lf[min(s + hi, cn - 1)]) + kw <= W + 15)
                        ++hi;
                    if (mc == KW) spread_step<R, EMAX, KW, KW>(Lk, s, lo, hi, lane, uoff, acc);
                    else if constexpr (KW > 1) {
                        if (mc == KW - 1) spread_step<R, EMAX, KW - 1, KW>(Lk, s, lo, hi, lane, uoff, acc);
                    }
                    lo = hi;
                }
            }
        }
        __syncthreads();
    }
    if (nchunk > 0) flush(W + 16);
}

// Block size and LDS bytes of the two schedules.
template <bool SPLIT> struct SpreadSched {
    static constexpr int threads = SPLIT ? 512 : 256;
    template <typename R, int EMAX> static constexpr size_t lds()
    {
        return (SPLIT ? 2 : 1) * sizeof(SpreadLds<R, EMAX>);
    }
};

template <typename R, int EMAX, bool SP, int KW>
__device__ __forceinline__ void spread_mfma_dispatch(int kw, const TfDesc &T, const AsyBoxD &bx,
                                                     const AsyItem &it, const SpreadRange &rg,
                                                     int nprod,
                                                     typename Cplx<R>::T *__restrict__ grid,
                                                     const double *__restrict__ ft,
                                                     SpreadLds<R, EMAX> *L)
{
    if constexpr (KW >= 1) {
        if (kw == KW) spread_mfma_body<R, EMAX, KW, SP>(T, bx, it, rg, nprod, grid, ft, L);
        else spread_mfma_dispatch<R, EMAX, SP, KW - 1>(kw, T, bx, it, rg, nprod, grid, ft, L);
    }
}

// (split schedule, EMAX <= 10: at most 128 VGPRs, so that two blocks share a
// CU; the few values that then leave registers are reloaded once per chunk)
template <typename R, int EMAX, bool SP>
__global__ __launch_bounds__(SpreadSched<SP>::threads, (SP && EMAX <= 10) ? 4 : 1) void k_asy_spread_mfma(
    const TfDesc *__restrict__ tfs, const AsyBoxD *__restrict__ boxes,
    const AsyItem *__restrict__ items, const SpreadRange *__restrict__ ranges,
    typename Cplx<R>::T *__restrict__ grid, const double *__restrict__ ft)
{
    // one or two SpreadLds buffers, dynamic (above 64 KiB for the widest EMAX)
    extern __shared__ unsigned char s_spread[];
    SpreadLds<R, EMAX> *L = (SpreadLds<R, EMAX> *)s_spread;
    const int bi = xcd_item<4>();
    const AsyItem it = items[bi];
    const SpreadRange rg = ranges[bi];
    const AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    const int nsub = 16 / it.tb;
    // products: 2 per exponent group; the per-wave count is uniform per
    // block, one branch-free body per count
    const int nprod = 2 * ((bx.E + nsub - 1) / nsub);
    const int kw = (nprod + 3) >> 2;
    constexpr int KWMAX = (2 * EMAX + 3) / 4;
    spread_mfma_dispatch<R, EMAX, SP, KWMAX>(kw, T, bx, it, rg, nprod, grid, ft, L);
}

// Interpolation, target-dense tiles: 16 consecutive targets whose kernel
// supports all fall in a window of nwin <= 16 grid positions starting at
// mw0 (item.pad holds mw0, item.tb the column chunk; the window length is
// derived on the device).  For exponent group g (nsub = 16/tb exponents):
//   D[i][(t, e2)] = sum_m phi(s_i - m Delta) grid[e][t][m],  e = g nsub + e2
// with the phi matrix computed once per tile; the FP64 epilogue applies
// e^{i(phi0 + w C)} h Delta / psihat(s) and the Hankel weights.
// (GR: the grid's precision; CR: the products' precision.)  The Hankel
// combination is folded into a running complex sum per (target, column):
// S = sum_e A'_e i^(e mod 2) G_e with A'_e = +-a_e (rho w)^-e, and the output
// is b0 Re(f S) with f = e^{i(phi0 + w C)} h Delta / psihat(s): the real part
// for even e and minus the imaginary part for odd e, as the reference's
// cos / sin terms (src/expansions.c:523-531, :549-557).
template <typename GR, typename CR, int EMAX>
__global__ __launch_bounds__(64) void k_asy_interp_mfma(const TfDesc *__restrict__ tfs,
                                                        const AsyBoxD *__restrict__ boxes,
                                                        const AsyItem *__restrict__ items,
                                                        const typename Cplx<GR>::T *__restrict__ grid,
                                                        const double *__restrict__ ft)
{
    typedef CR R;
    typedef typename Cplx<GR>::T C2;
    typedef typename Mfma<R>::V V;
    __shared__ double s_s[16], s_fr[16], s_fi[16], s_b0[16];
    __shared__ double s_P[16][EMAX];
    __shared__ double s_out[16][17];   // result tile, transposed for the atomic adds
    __shared__ int s_m0[16];
    AsyItem it = items[xcd_item<16>()];
    AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    const int lane = threadIdx.x;
    const int tb = it.tb, nsub = 16 / tb;
    const int E = bx.E, kw = T.kw;
    const int mw0 = it.pad;
    if (lane < 16) {
        int ti = lane;
        double w = T.w[bx.i0 + it.start + min(ti, it.count - 1)];
        double s = w - bx.D;
        s_s[ti] = s;
        s_m0[ti] = (int)ceil((s - bx.a2) / bx.delta);
        double nu = (double)T.order + 0.5 * (double)T.half;
        double sp, cp;
        sincos_cw(w * bx.C, &sp, &cp);
        int q = (int)lround(2.0 * nu + 1.0) & 7;
        double c0 = cos_qpi4(q), s0 = -cos_qpi4((q + 6) & 7);
        double amp = bx.h * bx.delta / (bx.a1 * ft_eval(ft + T.ft_off, s * bx.a1, T.kmax));
        s_fr[ti] = (cp * c0 - sp * s0) * amp;
        s_fi[ti] = (sp * c0 + cp * s0) * amp;
        s_b0[ti] = ti < it.count ? T.sign * sqrt(2.0 / kPi) / sqrt(w) : 0.0;
        double rw = 1.0 / (bx.rho * w), pw = 1.0;
        for (int e = 0; e < EMAX; ++e) {
            s_P[ti][e] = e < E ? ((((e >> 1) & 1) ? -T.a[e] : T.a[e]) * pw) : 0.0;
            pw *= rw;
        }
    }
    __syncthreads();
    int nwin = s_m0[it.count - 1] + kw - mw0;
    int nsteps = min(4, (nwin + 3) >> 2);
    // FP32 grids: the window starts at an even position (it still covers
    // every support: the tile test leaves one position of slack) and each
    // lane loads two adjacent complex values with one 16-byte load; K step
    // st then holds window positions 8 (st >> 1) + 2 ka + (st & 1).  Pairs
    // never straddle the end of the grid (Nf is even).  Positions outside a
    // target's support get phi = 0 from the kernel itself.
    constexpr bool PAIR = sizeof(C2) == 8;
    const int ia = lane & 15, ka = lane >> 4;
    const int mwb = PAIR ? (mw0 & ~1) : mw0;
    auto kpos = [&](int st) {
        return PAIR ? mwb + 8 * (st >> 1) + 2 * ka + (st & 1) : mwb + 4 * st + ka;
    };
    // A operand (phi) for each K step, reused by every exponent group
    R phiA[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        int m = kpos(st);
        double u = (s_s[ia] - (double)m * bx.delta) / bx.a2;
        phiA[st] = ((PAIR || st < nsteps) && ia < it.count) ? (R)es_kernel(u, T.beta) : (R)0;
    }
    const int jb = lane & 15, tbk = jb % tb, e2b = jb / tb;
    int moff[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        int m = PAIR ? kpos(st & 2) : kpos(st);     // PAIR: steps 2s, 2s+1 share one load
        moff[st] = m < 0 ? m + bx.Nf : (m >= bx.Nf ? m - bx.Nf : m);
    }
    const int G = (E + nsub - 1) / nsub;
    // every column chunk of the tile, with its phi and per-target factors
    for (int cb = it.col0; cb < T.B; cb += tb) {
    const int tcol = cb + tbk;
    const bool colok = tcol < T.B;
    double Sr[4] = {0.0, 0.0, 0.0, 0.0}, Si[4] = {0.0, 0.0, 0.0, 0.0};
    for (int g = 0; g < G; ++g) {
        const int eb = g * nsub + e2b;
        const bool eok = colok && eb < E;
        // the step's 4 grid values: loads at valid addresses, issued together
        const C2 *row = grid + bx.goff + (long long)((eok ? eb : 0) * T.B + (colok ? tcol : 0)) * bx.Nf;
        C2 b[4];
        if constexpr (PAIR) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float4 v = *(const float4 *)(row + moff[2 * h]);
                b[2 * h].x = v.x;
                b[2 * h].y = v.y;
                b[2 * h + 1].x = v.z;
                b[2 * h + 1].y = v.w;
            }
        } else {
#pragma unroll
            for (int st = 0; st < 4; ++st) b[st] = row[moff[st]];
        }
        V dr0, di0, dr1, di1;
#pragma unroll
        for (int v = 0; v < 4; ++v) { dr0[v] = (R)0; di0[v] = (R)0; dr1[v] = (R)0; di1[v] = (R)0; }
        const R zero = (R)0;
        dr0 = Mfma<R>::mma(phiA[0], eok ? (R)b[0].x : zero, dr0);
        di0 = Mfma<R>::mma(phiA[0], eok ? (R)b[0].y : zero, di0);
        dr1 = Mfma<R>::mma(phiA[1], eok ? (R)b[1].x : zero, dr1);
        di1 = Mfma<R>::mma(phiA[1], eok ? (R)b[1].y : zero, di1);
        dr0 = Mfma<R>::mma(phiA[2], eok ? (R)b[2].x : zero, dr0);
        di0 = Mfma<R>::mma(phiA[2], eok ? (R)b[2].y : zero, di0);
        dr1 = Mfma<R>::mma(phiA[3], eok ? (R)b[3].x : zero, dr1);
        di1 = Mfma<R>::mma(phiA[3], eok ? (R)b[3].y : zero, di1);
        const bool odd = eb & 1;
        const int ec = eok ? eb : 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int i = Mfma<R>::row(lane, v);
            double p = s_P[i][ec];
            double gr = (double)dr0[v] + (double)dr1[v], gi = (double)di0[v] + (double)di1[v];
            // S += p i^(e mod 2) G
            Sr[v] = fma(p, odd ? -gi : gr, Sr[v]);
            Si[v] = fma(p, odd ? gr : gi, Si[v]);
        }
    }
    double out[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        int i = Mfma<R>::row(lane, v);
        out[v] = s_b0[i] * (s_fr[i] * Sr[v] - s_fi[i] * Si[v]);
    }
    // fold the exponent sub-split (lanes j and j + tb * e2 share a column)
    for (int off = tb; off < 16; off <<= 1)
#pragma unroll
        for (int v = 0; v < 4; ++v) out[v] += __shfl_xor(out[v], off, 64);
    // through LDS, so that each atomic instruction adds 16 consecutive rows
    // of 4 columns (4 whole 128-byte lines of g) instead of 4 rows of 16
    if (e2b == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_out[Mfma<R>::row(lane, v)][jb] = out[v];
    }
    __syncthreads();
    for (int idx = lane; idx < 16 * tb; idx += 64) {
        const int c = idx >> 4, i = idx & 15;
        if (i < it.count && cb + c < T.B)
            atomicAdd(&T.gs[(long long)(cb + c) * T.ld_gs + bx.i0 + it.start + i], s_out[i][c]);
    }
    __syncthreads();
    }
}

// Interpolation, deconvolution, phase and the reference's output weights
// (src/expansions.c:523-531, :549-557).  Per-target factors (kernel weights
// phi, e^{i (phi0 + w C)} h Delta / psihat(s), weight base) are computed in
// FP64 by all lanes of the block together and shared through LDS; the grid
// window a tile of (sorted) targets touches is staged in LDS for a group of
// exponents at a time.  Lanes run over targets fastest, so lanes of one
// column read neighbouring window entries.
constexpr int kInterpLds = 20 * 1024;   // bytes of staged grid window

// Offsets (in doubles) of the interpolation kernel's dynamic LDS regions:
// part 0 = per-target scalars after phi[tmax][16]; part 1 = the grid window,
// 16-byte aligned; part 2 = total size.
__host__ __device__ inline size_t interp_lds_off(int tmax, int part, size_t rsize)
{
    size_t phi = ((size_t)tmax * 16 * rsize + 15) / 16 * 2;
    if (part == 0) return phi;
    size_t win = phi + 4 * (size_t)tmax + ((size_t)tmax + 1) / 2;
    win = (win + 1) / 2 * 2;
    if (part == 1) return win;
    return win + kInterpLds / 8;
}

template <typename R, int EMAX>
__global__ __launch_bounds__(256) void k_asy_interp(const TfDesc *__restrict__ tfs,
                                                    const AsyBoxD *__restrict__ boxes,
                                                    const AsyItem *__restrict__ items,
                                                    const typename Cplx<R>::T *__restrict__ grid,
                                                    const double *__restrict__ ft, int tmax)
{
    typedef typename Cplx<R>::T C2;
    // dynamic LDS, sized per launch for the widest target tile (tmax):
    // phi[tmax][16] | fr, fi, b0, rw [tmax] | m0[tmax] | grid window
    extern __shared__ double s_dyn[];
    AsyItem it = items[xcd_item<4>()];
    AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    int tb = it.tb, ti_n = 256 / tb;
    R(*s_phi)[16] = (R(*)[16])s_dyn;
    double *s_fr = s_dyn + interp_lds_off(tmax, 0, sizeof(R));
    double *s_fi = s_fr + tmax, *s_b0 = s_fi + tmax, *s_rw = s_b0 + tmax;
    int *s_m0 = (int *)(s_rw + tmax);
    C2 *s_g = (C2 *)(s_dyn + interp_lds_off(tmax, 1, sizeof(R)));
    int E = bx.E, kw = T.kw;
    int nt = min(it.count, ti_n);
    double inv_a2 = 1.0 / bx.a2;
    // per-target scalars: one thread per target
    for (int ti = threadIdx.x; ti < nt; ti += 256) {
        double w = T.w[bx.i0 + it.start + ti];
        double s = w - bx.D;
        s_m0[ti] = (int)ceil((s - bx.a2) / bx.delta);
        double nu = (double)T.order + 0.5 * (double)T.half;
        double sp, cp;
        sincos(w * bx.C, &sp, &cp);
        int q = (int)lround(2.0 * nu + 1.0) & 7;   // phi0 = -q pi / 4
        double c0 = cos_qpi4(q), s0 = -cos_qpi4((q + 6) & 7);
        double amp = bx.h * bx.delta / (bx.a1 * ft_eval(ft + T.ft_off, s * bx.a1, T.kmax));
        s_fr[ti] = (cp * c0 - sp * s0) * amp;
        s_fi[ti] = (sp * c0 + cp * s0) * amp;
        s_b0[ti] = T.sign * sqrt(2.0 / kPi) / sqrt(w);
        s_rw[ti] = 1.0 / (bx.rho * w);
    }
    __syncthreads();
    // kernel weights: (target, tap) pairs over all lanes
    for (int q = threadIdx.x; q < nt * kw; q += 256) {
        int ti = q / kw, k = q % kw;
        double s = T.w[bx.i0 + it.start + ti] - bx.D;
        s_phi[ti][k] = (R)es_kernel((s - (double)(s_m0[ti] + k) * bx.delta) * inv_a2, T.beta);
    }
    int pw0 = s_m0[0];
    int pw = s_m0[nt - 1] + kw - pw0;
    int ge = kInterpLds / (int)sizeof(C2) / max(1, pw * tb);   // exponents staged per pass
    bool use_lds = ge >= 1;
    if (ge > E) ge = E;
    __syncthreads();
    int ti = threadIdx.x % ti_n, tl = threadIdx.x / ti_n;
    int t = it.col0 + tl;
    bool active = ti < it.count && t < T.B;
    int m0 = active ? s_m0[ti] : 0;
    double fr = active ? s_fr[ti] : 0.0, fi = active ? s_fi[ti] : 0.0;
    double base = active ? s_b0[ti] : 0.0, rw = active ? s_rw[ti] : 0.0;
    R phi[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) phi[k] = (active && k < kw) ? s_phi[ti][k] : (R)0;
    double Sr = 0.0, Si = 0.0;
    if (use_lds) {
        for (int e0 = 0; e0 < E; e0 += ge) {
            int ne = min(ge, E - e0);
            // the window rows of ne exponents, 8 loads per thread in flight at a time
            for (int q0 = 0; q0 < ne * tb * pw; q0 += 256 * 8) {
                C2 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    int q = q0 + threadIdx.x + 256 * u;
                    int pos = q % pw, row = q / pw;
                    int ee = row / tb, tt = row % tb;
                    int tc = it.col0 + tt;
                    int m = pw0 + pos;
                    m = m < 0 ? m + bx.Nf : m;
                    v[u].x = (R)0;
                    v[u].y = (R)0;
                    if (q < ne * tb * pw && tc < T.B)
                        v[u] = grid[bx.goff + (long long)((e0 + ee) * T.B + tc) * bx.Nf + m];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    int q = q0 + threadIdx.x + 256 * u;
                    if (q < ne * tb * pw) s_g[q] = v[u];
                }
            }
            __syncthreads();
            if (active) {
                for (int ee = 0; ee < ne; ++ee) {
                    const C2 *row = s_g + (ee * tb + tl) * pw + (m0 - pw0);
                    R gr = (R)0, gi = (R)0;
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        if (k < kw) {
                            C2 v = row[k];
                            gr = fma(phi[k], v.x, gr);
                            gi = fma(phi[k], v.y, gi);
                        }
                    }
                    int e = e0 + ee;
                    double A = (((e >> 1) & 1) ? -T.a[e] : T.a[e]) * base;
                    // S += A i^(e mod 2) G; the output is Re(f S)
                    Sr = fma(A, (e & 1) ? -(double)gi : (double)gr, Sr);
                    Si = fma(A, (e & 1) ? (double)gr : (double)gi, Si);
                    base *= rw;
                }
            }
            __syncthreads();
        }
    } else if (active) {
        for (int e = 0; e < E; ++e) {
            const C2 *row = grid + bx.goff + (long long)(e * T.B + t) * bx.Nf;
            R gr = (R)0, gi = (R)0;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                if (k < kw) {
                    int m = m0 + k;
                    m = m < 0 ? m + bx.Nf : m;
                    C2 v = row[m];
                    gr = fma(phi[k], v.x, gr);
                    gi = fma(phi[k], v.y, gi);
                }
            }
            double A = (((e >> 1) & 1) ? -T.a[e] : T.a[e]) * base;
            Sr = fma(A, (e & 1) ? -(double)gi : (double)gr, Sr);
            Si = fma(A, (e & 1) ? (double)gr : (double)gi, Si);
            base *= rw;
        }
    }
    if (active)
        atomicAdd(&T.gs[(long long)t * T.ld_gs + bx.i0 + it.start + ti], fr * Sr - fi * Si);
}

// ------------------------------------------------------------------------
// Row permutations (callers' unsorted grids) and gradient combination.
__global__ void k_gather_rows(const double *__restrict__ src, long long ld_src,
                              double *__restrict__ dst, long long ld_dst,
                              const int *__restrict__ perm, int rows, int cols)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)rows * cols) return;
    int i = (int)(q % rows), t = (int)(q / rows);
    dst[(long long)t * ld_dst + i] = src[(long long)t * ld_src + (perm ? perm[i] : i)];
}

__global__ void k_scatter_rows(const double *__restrict__ src, long long ld_src,
                               double *__restrict__ dst, long long ld_dst,
                               const int *__restrict__ perm, int rows, int cols)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)rows * cols) return;
    int i = (int)(q % rows), t = (int)(q / rows);
    dst[(long long)t * ld_dst + (perm ? perm[i] : i)] = src[(long long)t * ld_src + i];
}

// out[k,t] = in[k,t] * x[k] (column scaling by a node array, sorted order)
__global__ void k_scale_rows(const double *__restrict__ in, long long ld_in,
                             double *__restrict__ out, long long ld_out,
                             const double *__restrict__ x, int rows, int cols)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)rows * cols) return;
    int i = (int)(q % rows), t = (int)(q / rows);
    out[(long long)t * ld_out + i] = in[(long long)t * ld_in + i] * x[i];
}

// grad[k] = sum_t a[k,t] * ( (nu / x_k) * P[k,t] - sgn * Q[k,t] ), sorted rows;
// x_k == 0 uses the limit of (nu/x) J_nu(x w) sum: lim0[t] for nu == 1.
__global__ void k_grad_combine(const double *__restrict__ a, long long ld_a,
                               const double *__restrict__ P, long long ld_p,
                               const double *__restrict__ Q, long long ld_q,
                               const double *__restrict__ x, const double *__restrict__ lim0,
                               double nu, double sgn, int rows, int cols,
                               double *__restrict__ grad)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rows) return;
    double xk = x[k];
    double acc = 0.0;
    for (int t = 0; t < cols; ++t) {
        double pterm;
        if (xk != 0.0) pterm = (nu / xk) * P[(long long)t * ld_p + k];
        else pterm = (lim0 && nu == 1.0) ? sgn * 0.5 * lim0[t] : (nu == 0.5 ? INFINITY : 0.0);
        acc = fma(a[(long long)t * ld_a + k], pterm - sgn * Q[(long long)t * ld_q + k], acc);
    }
    grad[k] = acc;
}

// lim0[t] = sum_i rho[i,t] * x[i]
__global__ void k_colsum_weighted(const double *__restrict__ rho, long long ld, const double *__restrict__ x,
                                  int rows, int cols, double *__restrict__ out)
{
    __shared__ double red[256];
    int t = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < rows; i += 256) acc += rho[(long long)t * ld + i] * x[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = red[0];
}

}  // namespace nufht_dev
