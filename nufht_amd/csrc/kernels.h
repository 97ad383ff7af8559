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
    int pad0;
    double beta;           // ES shape parameter
    double kmax;           // domain of the kernel-transform table
    double sign;           // -1 for negative odd integer orders
    double a[kMaxA];       // Hankel coefficients a_k(|nu|)
};

struct DirItem { int tf, i0, ni, j0, nj, tpr, col0, pad; };
struct LocBoxD { int tf, i0, ni, j0, nj, coef_off; double r_end; };
struct LocItem { int box, start, count, col0; };
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

__global__ __launch_bounds__(256) void k_dir(const TfDesc *__restrict__ tfs,
                                             const DirItem *__restrict__ items)
{
    __shared__ double s_r[kDirTJ];
    __shared__ double s_c[kDirTJ * kDirBM];
    DirItem it = items[blockIdx.x];
    const TfDesc &T = tfs[it.tf];
    int nb = min(kDirBM, T.B - it.col0);
    for (int q = threadIdx.x; q < it.nj; q += 256) s_r[q] = T.r[it.j0 + q];
    for (int q = threadIdx.x; q < it.nj * kDirBM; q += 256) {
        int j = q % it.nj, t = q / it.nj;
        s_c[j * kDirBM + t] = t < nb ? T.cs[(long long)(it.col0 + t) * T.ld_cs + it.j0 + j] : 0.0;
    }
    __syncthreads();
    int tpr = it.tpr;
    int row = threadIdx.x / tpr, lane = threadIdx.x % tpr;
    bool active = row < it.ni;
    double acc[kDirBM];
#pragma unroll
    for (int t = 0; t < kDirBM; ++t) acc[t] = 0.0;
    if (active) {
        double w = T.w[it.i0 + row];
        for (int j = lane; j < it.nj; j += tpr) {
            double J = besselj_wr(T.order, T.half, w, s_r[j]);
#pragma unroll
            for (int t = 0; t < kDirBM; ++t) acc[t] = fma(J, s_c[j * kDirBM + t], acc[t]);
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
// Local expansion, projection half: coef[l][t] += sum_j B_l(x_j) c_j,t with
// x_j = r_j / r_end.  Integer orders follow the reference's Wimp expansion
// (src/expansions.c:242-273): even nu -> eps_l T_2l(x), odd nu -> 2 T_2l+1(x).
// Half-integer orders (Chebyshev interpolation of j_n, this build) use
// T_2l+p(x), p = n mod 2, with weights c_j sqrt(r_j).
// Chebyshev polynomials come from the three-term recurrence.
constexpr int kLocBM = 16;

__global__ __launch_bounds__(256) void k_loc_proj(const TfDesc *__restrict__ tfs,
                                                  const LocBoxD *__restrict__ boxes,
                                                  const LocItem *__restrict__ items,
                                                  double *__restrict__ coef, int ch)
{
    extern __shared__ double sm[];
    LocItem it = items[blockIdx.x];
    LocBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    int K = T.K_loc;
    double *sT = sm;                       // [(K+1)][ch]
    double *sC = sm + (size_t)(K + 1) * ch; // [ch][kLocBM]
    int nb = min(kLocBM, T.B - it.col0);
    int par = T.half ? (T.order & 1) : (T.order & 1);
    for (int q = threadIdx.x; q < it.count; q += blockDim.x) {
        int j = bx.j0 + it.start + q;
        double rj = T.r[j];
        double x = rj / bx.r_end;
        x = fmin(1.0, fmax(-1.0, x));
        double tm1 = 1.0, t0 = x;          // T_0, T_1
        // basis index l <-> Chebyshev degree 2l + par
        if (par == 0) sT[q] = 1.0;
        else sT[q] = 2.0 * x;
        int deg = 1;
        for (int l = 1; l <= K; ++l) {
            int target = 2 * l + par;
            while (deg < target) {
                double tn = 2.0 * x * t0 - tm1;
                tm1 = t0;
                t0 = tn;
                deg++;
            }
            sT[(size_t)l * ch + q] = 2.0 * t0;
        }
        double wgt = T.half ? sqrt(rj) : 1.0;
        for (int t = 0; t < kLocBM; ++t)
            sC[q * kLocBM + t] = t < nb ? wgt * T.cs[(long long)(it.col0 + t) * T.ld_cs + j] : 0.0;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < (K + 1) * nb; p += blockDim.x) {
        int l = p / nb, t = p % nb;
        double acc = 0.0;
        const double *tl = sT + (size_t)l * ch;
        for (int q = 0; q < it.count; ++q) acc = fma(tl[q], sC[q * kLocBM + t], acc);
        atomicAdd(&coef[bx.coef_off + (long long)l * T.B + it.col0 + t], acc);
    }
}

// Local expansion, evaluation half (reference src/expansions.c:275-336):
// integer orders compute J_0..J_{l0+K+odd}(z), z = w r_end / 2, by one
// Miller sweep and form b_l = J_{l0+odd+l}(z) J_{|l0-l|}(z) (sign (-1) for
// l > l0 with l - l0 odd); half-integer orders sample j_n(w r_end x_q) at
// the K+1 positive Chebyshev nodes and map them to coefficients with the
// precomputed matrix cheb[].  Then g[i,t] += sum_l b_l coef[l][t].
__global__ void k_loc_eval(const TfDesc *__restrict__ tfs, const LocBoxD *__restrict__ boxes,
                           const LocItem *__restrict__ items, const double *__restrict__ coef,
                           const double *__restrict__ cheb, int stride_j)
{
    extern __shared__ double sm[];
    LocItem it = items[blockIdx.x];
    LocBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    int K = T.K_loc;
    int tid = threadIdx.x, nthr = blockDim.x;
    double *sJ = sm + (size_t)tid * stride_j;              // per-thread scratch
    double *sCoef = sm + (size_t)nthr * stride_j;          // [(K+1)][kLocBM]
    int nb = min(kLocBM, T.B - it.col0);
    for (int p = tid; p < (K + 1) * kLocBM; p += nthr) {
        int l = p / kLocBM, t = p % kLocBM;
        sCoef[p] = t < nb ? coef[bx.coef_off + (long long)l * T.B + it.col0 + t] : 0.0;
    }
    __syncthreads();
    if (tid >= it.count) return;
    int i = bx.i0 + it.start + tid;
    double w = T.w[i];
    double scale = 1.0;
    double *b = sJ;   // b_l overwrites the front of the scratch once used
    if (!T.half) {
        int l0 = T.order >> 1, odd = T.order & 1;
        int maxo = l0 + K + odd;
        double z = w * bx.r_end * 0.5;
        double *J = sJ + (K + 1);
        if (z == 0.0) {
            for (int k = 0; k <= maxo; ++k) J[k] = (k == 0) ? 1.0 : 0.0;
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
                if (km1 <= maxo) J[km1] = j;
                if (km1 > 0 && !(km1 & 1)) sum += 2.0 * j;
                if (fabs(j) > 1e200) {
                    j *= 1e-200; jp1 *= 1e-200; sum *= 1e-200;
                    for (int q = km1; q <= maxo; ++q) J[q] *= 1e-200;
                }
            }
            double inv = 1.0 / (sum + j);
            for (int k = 0; k <= maxo; ++k) J[k] *= inv;
        }
        for (int l = 0; l <= K; ++l) {
            int lo = l0 - l;
            lo = lo < 0 ? -lo : lo;
            double v = J[l0 + odd + l] * J[lo];
            if (l > l0 && ((l - l0) & 1)) v = -v;
            b[l] = v;
        }
    } else {
        // f_q = j_n(z x_q), x_q = cos(pi (q + 1/2) / (2K + 2)), q = 0..K
        double z = w * bx.r_end;
        double *f = sJ + (K + 1);
        for (int q = 0; q <= K; ++q) {
            double xq = cos(kPi * ((double)q + 0.5) / (double)(2 * K + 2));
            double hi = z * xq, lo = fma(z, xq, -hi);
            f[q] = sph_jn(T.order, hi, lo);
        }
        const double *M = cheb + T.cheb_off;
        for (int l = 0; l <= K; ++l) {
            double a = 0.0;
            for (int q = 0; q <= K; ++q) a = fma(M[l * (K + 1) + q], f[q], a);
            b[l] = a;
        }
        scale = sqrt(2.0 * w / kPi);
    }
    for (int t = 0; t < nb; ++t) {
        double acc = 0.0;
        for (int l = 0; l <= K; ++l) acc = fma(b[l], sCoef[l * kLocBM + t], acc);
        atomicAdd(&T.gs[(long long)(it.col0 + t) * T.ld_gs + i], T.sign * scale * acc);
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
// positions, a column chunk and a source sub-range [jlo, jhi); items that
// split a tile's sources (pad == 1) add atomically into a zeroed tile.
template <typename R, int EMAX>
__global__ __launch_bounds__(256) void k_asy_spread(const TfDesc *__restrict__ tfs,
                                                    const AsyBoxD *__restrict__ boxes,
                                                    const AsyItem *__restrict__ items,
                                                    const int2 *__restrict__ ranges,
                                                    typename Cplx<R>::T *__restrict__ grid,
                                                    const double *__restrict__ ft)
{
    __shared__ R s_zr[kSpreadQ][kTapMax], s_zi[kSpreadQ][kTapMax];
    __shared__ R s_c[kSpreadQ][16];
    __shared__ R s_ri[kSpreadQ];
    __shared__ int s_pf[kSpreadQ];
    AsyItem it = items[blockIdx.x];
    int2 jr = ranges[blockIdx.x];
    AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    int tb = it.tb;
    int tl = threadIdx.x % tb, pl = threadIdx.x / tb;
    int p = it.start + pl;
    int l = p <= bx.L ? p : p - bx.Nf;
    bool valid = (pl < it.count) && (l >= -bx.L) && (l <= bx.L);
    int t = it.col0 + tl;
    bool active = valid && (t < T.B);
    int kw = T.kw;
    double beta = T.beta;
    double inv_a1 = 1.0 / bx.a1;
    R acc_r[EMAX], acc_i[EMAX];
#pragma unroll
    for (int e = 0; e < EMAX; ++e) { acc_r[e] = (R)0; acc_i[e] = (R)0; }
    int E = bx.E;
    for (int c0 = jr.x; c0 < jr.y; c0 += kSpreadQ) {
        int cn = min(kSpreadQ, jr.y - c0);
        for (int q = threadIdx.x; q < cn; q += 256) {
            double rj = T.r[c0 + q];
            double x = rj - bx.C;
            double sn, cs;
            sincos(bx.D * x, &sn, &cs);
            double rs = 1.0 / sqrt(rj);
            double br = cs * rs, bi = sn * rs;
            int lf = (int)ceil((x - bx.a1) / bx.h);
            s_pf[q] = lf;
            s_ri[q] = (R)(bx.rho / rj);
            for (int k = 0; k < kTapMax; ++k) {
                double psi = 0.0;
                if (k <= kw) psi = (double)es_eval((R)(((double)(lf + k) * bx.h - x) * inv_a1), (R)beta);
                s_zr[q][k] = (R)(psi * br);
                s_zi[q][k] = (R)(psi * bi);
            }
        }
        for (int q = threadIdx.x; q < cn * tb; q += 256) {
            int jq = q / tb, tt = q % tb;
            int tc = it.col0 + tt;
            s_c[jq][tt] = tc < T.B ? (R)T.cs[(long long)tc * T.ld_cs + c0 + jq] : (R)0;
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
                int k = l - s_pf[q];
                R cv = s_c[q][tl];
                R zr = s_zr[q][k] * cv, zi = s_zi[q][k] * cv;
                R ri = s_ri[q];
#pragma unroll
                for (int e = 0; e < EMAX; ++e) {
                    if (e < E) {
                        acc_r[e] += zr;
                        acc_i[e] += zi;
                        zr *= ri;
                        zi *= ri;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (pl < it.count && t < T.B) {
        double dec = 0.0;
        if (valid) dec = 1.0 / (bx.a2 * ft_eval(ft + T.ft_off, (double)l * bx.h * bx.a2, T.kmax));
        for (int e = 0; e < E; ++e) {
            typename Cplx<R>::T *dst = grid + bx.goff + ((long long)(e * T.B + t)) * bx.Nf + p;
            R vr = (R)((double)acc_r[e] * dec), vi = (R)((double)acc_i[e] * dec);
            if (it.pad) {
                if (valid) grid_add(dst, vr, vi);
            } else {
                typename Cplx<R>::T v;
                v.x = vr;
                v.y = vi;
                *dst = v;
            }
        }
    }
}

// Interpolation, deconvolution, phase and the reference's output weights
// (src/expansions.c:523-531, :549-557).  Kernel weights phi, the factor
// e^{i (phi0 + w C)} h Delta / psihat(s) and the weight base are computed
// once per target in FP64 and shared through LDS by the column lanes; the
// grid window a tile of (sorted) targets touches is staged in LDS per
// exponent when it is at most kInterpWin positions wide.
template <typename R, int EMAX>
__global__ __launch_bounds__(256) void k_asy_interp(const TfDesc *__restrict__ tfs,
                                                    const AsyBoxD *__restrict__ boxes,
                                                    const AsyItem *__restrict__ items,
                                                    const typename Cplx<R>::T *__restrict__ grid,
                                                    const double *__restrict__ ft)
{
    __shared__ R s_phi[256][16];
    __shared__ int s_m0[256];
    __shared__ double s_fr[256], s_fi[256], s_b0[256], s_rw[256];
    __shared__ typename Cplx<R>::T s_g[16][kInterpWin];
    AsyItem it = items[blockIdx.x];
    AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    int tb = it.tb, ti_n = 256 / tb;
    int E = bx.E, kw = T.kw;
    if (threadIdx.x < ti_n && threadIdx.x < it.count) {
        int ti = threadIdx.x;
        int i = bx.i0 + it.start + ti;
        double w = T.w[i];
        double s = w - bx.D;
        int m0 = (int)ceil((s - bx.a2) / bx.delta);
        double inv_a2 = 1.0 / bx.a2;
        for (int k = 0; k < kw; ++k)
            s_phi[ti][k] = (R)es_kernel((s - (double)(m0 + k) * bx.delta) * inv_a2, T.beta);
        s_m0[ti] = m0;
        // e^{i (phi0 + w C)}, phi0 = -pi (nu/2 + 1/4) = -q pi / 4
        double nu = (double)T.order + 0.5 * (double)T.half;
        double sp, cp;
        sincos(w * bx.C, &sp, &cp);
        int q = (int)lround(2.0 * nu + 1.0) & 7;
        double c0 = cos_qpi4(q), s0 = -cos_qpi4((q + 6) & 7);
        double fr = cp * c0 - sp * s0, fi = sp * c0 + cp * s0;
        double amp = bx.h * bx.delta / (bx.a1 * ft_eval(ft + T.ft_off, s * bx.a1, T.kmax));
        s_fr[ti] = fr * amp;
        s_fi[ti] = fi * amp;
        s_b0[ti] = T.sign * sqrt(2.0 / kPi) / sqrt(w);
        s_rw[ti] = 1.0 / (bx.rho * w);
    }
    __syncthreads();
    int nvalid = min(it.count, ti_n);
    int pw0 = s_m0[0];
    int pw = s_m0[nvalid - 1] + kw - pw0;
    bool use_lds = pw <= kInterpWin;
    int ti = threadIdx.x % ti_n, tl = threadIdx.x / ti_n;
    int t = it.col0 + tl;
    bool active = ti < it.count && t < T.B;
    int m0 = active ? s_m0[ti] : 0;
    double fr = active ? s_fr[ti] : 0.0, fi = active ? s_fi[ti] : 0.0;
    double base = active ? s_b0[ti] : 0.0, rw = active ? s_rw[ti] : 0.0;
    double out = 0.0;
    for (int e = 0; e < E; ++e) {
        if (use_lds) {
            for (int q = threadIdx.x; q < pw * tb; q += 256) {
                int tt = q / pw, pos = q % pw;
                int tc = it.col0 + tt;
                int m = pw0 + pos;
                m = m < 0 ? m + bx.Nf : m;
                if (tc < T.B) s_g[tt][pos] = grid[bx.goff + (long long)(e * T.B + tc) * bx.Nf + m];
            }
            __syncthreads();
        }
        if (active) {
            R gr = (R)0, gi = (R)0;
            if (use_lds) {
                const typename Cplx<R>::T *row = &s_g[tl][m0 - pw0];
                for (int k = 0; k < kw; ++k) {
                    R f = s_phi[ti][k];
                    gr += f * row[k].x;
                    gi += f * row[k].y;
                }
            } else {
                const typename Cplx<R>::T *row = grid + bx.goff + (long long)(e * T.B + t) * bx.Nf;
                for (int k = 0; k < kw; ++k) {
                    int m = m0 + k;
                    m = m < 0 ? m + bx.Nf : m;
                    typename Cplx<R>::T v = row[m];
                    R f = s_phi[ti][k];
                    gr += f * v.x;
                    gi += f * v.y;
                }
            }
            double vr = fr * (double)gr - fi * (double)gi, vi = fr * (double)gi + fi * (double)gr;
            double A = (((e >> 1) & 1) ? -T.a[e] : T.a[e]) * base;
            out = fma(A, (e & 1) ? -vi : vr, out);
            base *= rw;
        }
        if (use_lds) __syncthreads();
    }
    if (active) {
        int i = bx.i0 + it.start + ti;
        atomicAdd(&T.gs[(long long)t * T.ld_gs + i], out);
    }
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
