// Near-field direct sums, local expansions, zeroing, row permutations and
// gradient combination kernels (launched by engine.hip).
#pragma once
#include "device_common.h"

namespace nufht_dev {

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

// J_nu on [0, jtab_n): on unit interval iv a polynomial of degree kJDeg in
// u = x - (iv + 1/2), |u| <= 1/2, held as its monomial coefficients (Horner:
// one FMA per degree, where the Chebyshev form's Clenshaw recurrence takes
// two operations).  The polynomial is the interval's Chebyshev interpolant at
// kJDeg + 1 nodes, its basis changed exactly.  J_nu (integer nu) and
// J_nu(x) / sqrt(x) (half-integer nu, tabulated so: it is entire) are
// band-limited to frequencies |s| <= 1, so the Chebyshev coefficients on a
// unit interval fall as 2 (1/4)^n / n! and the monomial form is well
// conditioned (its values agree with the Chebyshev form's to a few ulps).
// Degree 15 keeps relative accuracy where J_nu is tiny (x < nu: J_10 on
// [0, 1] is below 3e-10), which degree 11 would lose.
constexpr int kJDeg = 15;

__device__ __forceinline__ double jtab_horner(const double (&c)[kJDeg + 1], double u)
{
    double p = c[kJDeg];
#pragma unroll
    for (int n = kJDeg - 1; n >= 0; --n) p = fma(p, u, c[n]);
    return p;
}

// Table build: one block of kJDeg + 1 threads per unit interval; values of
// the exact routine at Chebyshev nodes, the discrete cosine transform
// (Chebyshev coefficients in t = 2u), then the exact change of basis to
// monomials in u: T_k(2u) = sum_n M[k][n] u^n.
__global__ void k_jtab_build(int order, int half, int nint, double *__restrict__ tab)
{
    constexpr int N = kJDeg + 1;
    __shared__ double f[N], ch[N], M[N][N];
    int iv = blockIdx.x, k = threadIdx.x;
    double tk = cos(kPi * ((double)k + 0.5) / N);
    double x = (double)iv + 0.5 * (tk + 1.0);
    double v = half ? jn_half(order, x, 0.0) : jn_int(order, x, 0.0);
    if (half) v = x > 0.0 ? v / sqrt(x) : 0.0;
    f[k] = v;
    if (k == 0) {
        // T_0 = 1, T_1 = 2u, T_{j+1} = 4u T_j - T_{j-1} (integers, exact)
        for (int j = 0; j < N; ++j)
            for (int n = 0; n < N; ++n) M[j][n] = 0.0;
        M[0][0] = 1.0;
        M[1][1] = 2.0;
        for (int j = 1; j + 1 < N; ++j)
            for (int n = 0; n < N; ++n)
                M[j + 1][n] = (n > 0 ? 4.0 * M[j][n - 1] : 0.0) - M[j - 1][n];
    }
    __syncthreads();
    double acc = 0.0;
    for (int q = 0; q < N; ++q) acc += f[q] * cos(kPi * (double)k * ((double)q + 0.5) / N);
    ch[k] = acc * (k == 0 ? 1.0 : 2.0) / N;
    __syncthreads();
    double a = 0.0;
    for (int j = N - 1; j >= k; --j) a += ch[j] * M[j][k];
    tab[(long long)iv * N + k] = a;
}

// NB: the item's columns, rounded up to 4, 8 or 16 (one kernel per class).
// The tpr lanes of a row each take a contiguous, odd-length share of the
// staged sources (odd: the lanes' reads fall in different LDS banks); with
// sources sorted by r, x = k r rises slowly along a share, so the lane keeps
// the Chebyshev coefficients of its current unit interval in registers and
// reloads them only when x enters the next one.
template <int NB>
__global__ __launch_bounds__(256) void k_dir(const TfDesc *__restrict__ tfs,
                                             const DirItem *__restrict__ items,
                                             const int2 *__restrict__ rng,
                                             const double *__restrict__ jtab)
{
    // rows of NB + 2 doubles: 16-byte aligned for paired reads, and the odd
    // shares of the row's lanes still fall in different banks
    constexpr int LDC = NB + 2;
    __shared__ double s_r[kDirTJ];
    __shared__ __align__(16) double s_c[kDirTJ * LDC];
    DirItem it = items[blockIdx.x];
    const TfDesc &T = tfs[it.tf];
    int nb = min(NB, T.B - it.col0);
    int tpr = it.tpr;
    int row = threadIdx.x / tpr, lane = threadIdx.x % tpr;
    bool active = row < it.ni;
    double acc[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) acc[t] = 0.0;
    const double *tab = jtab + T.jtab_off;
    double xmax = (double)T.jtab_n;
    const double w = active ? ld_global(&T.w[it.i0 + row]) : 0.0;
    int civ = -1;                   // unit interval of the coefficients held in cc
    double cc[kJDeg + 1];
#pragma unroll
    for (int n = 0; n <= kJDeg; ++n) cc[n] = 0.0;
    for (int q = 0; q < it.rcnt; ++q) {
        const int2 rq = rng[it.roff + q];
        for (int c = 0; c < rq.y; c += kDirTJ) {
            const int nj = min(kDirTJ, rq.y - c), j0 = rq.x + c;
            __syncthreads();   // the previous chunk's readers are done
            {
                // one source per thread (nj <= kDirTJ = blockDim), its columns loaded together
                const int j = threadIdx.x;
                const bool ok = j < nj;
                double cv[NB];
#pragma unroll
                for (int t = 0; t < NB; ++t)
                    cv[t] = (ok && t < nb) ? ld_global(&T.cs[(long long)(it.col0 + t) * T.ld_cs + j0 + j]) : 0.0;
                s_r[j] = ok ? ld_global(&T.r[j0 + j]) : 0.0;
#pragma unroll
                for (int t = 0; t < NB; t += 2) *(double2 *)&s_c[j * LDC + t] = make_double2(cv[t], cv[t + 1]);
            }
            __syncthreads();
            if (active) {
                const int per = ((nj + tpr - 1) / tpr) | 1;
                const int ja = lane * per, jb = min(nj, ja + per);
                auto load_iv = [&](int iv) {
                    if (iv != civ) {
                        civ = iv;
                        const double *cf = tab + (long long)iv * (kJDeg + 1);
#pragma unroll
                        for (int n = 0; n <= kJDeg; ++n) cc[n] = cf[n];
                    }
                };
                // J at x = w r_j: the table (its interval's coefficients in
                // registers), or the exact routine beyond it
                auto one = [&](int j) {
                    const double x = w * s_r[j];
                    if (!(x >= 0.0 && x < xmax)) return besselj_wr(T.order, T.half, w, s_r[j]);
                    const int iv = (int)x;
                    load_iv(iv);
                    const double J = jtab_horner(cc, x - ((double)iv + 0.5));
                    return T.half ? J * sqrt(x) : J;
                };
                int j = ja;
                // 16 columns: two sources at a time; in one interval (the
                // common case) their Horner chains run interleaved
                // (for fewer columns the second chain's registers cost more
                // occupancy than the interleaving gains)
                for (; NB == 16 && j + 1 < jb; j += 2) {
                    const double x0 = w * s_r[j], x1 = w * s_r[j + 1];
                    double J0, J1;
                    const int iv0 = (int)x0;
                    if (x0 >= 0.0 && x1 < xmax && iv0 == (int)x1) {
                        load_iv(iv0);
                        const double mid = (double)iv0 + 0.5;
                        J0 = jtab_horner(cc, x0 - mid);
                        J1 = jtab_horner(cc, x1 - mid);
                        if (T.half) {
                            J0 *= sqrt(x0);
                            J1 *= sqrt(x1);
                        }
                    } else {
                        J0 = one(j);
                        J1 = one(j + 1);
                    }
                    const double2 *cr0 = (const double2 *)&s_c[j * LDC];
                    const double2 *cr1 = (const double2 *)&s_c[(j + 1) * LDC];
#pragma unroll
                    for (int t = 0; t < NB / 2; ++t) {
                        const double2 v = cr0[t], u = cr1[t];
                        acc[2 * t] = fma(J1, u.x, fma(J0, v.x, acc[2 * t]));
                        acc[2 * t + 1] = fma(J1, u.y, fma(J0, v.y, acc[2 * t + 1]));
                    }
                }
                for (; j < jb; ++j) {
                    const double J = one(j);
                    const double2 *cr = (const double2 *)&s_c[j * LDC];
#pragma unroll
                    for (int t = 0; t < NB / 2; ++t) {
                        const double2 v = cr[t];
                        acc[2 * t] = fma(J, v.x, acc[2 * t]);
                        acc[2 * t + 1] = fma(J, v.y, acc[2 * t + 1]);
                    }
                }
            }
        }
    }
    for (int off = tpr >> 1; off > 0; off >>= 1) {
#pragma unroll
        for (int t = 0; t < NB; ++t) acc[t] += __shfl_xor(acc[t], off, 64);
    }
    if (active && lane == 0) {
        for (int t = 0; t < nb; ++t)
            atomic_add_global(&T.gs[(long long)(it.col0 + t) * T.ld_gs + it.i0 + row], T.sign * acc[t]);
    }
}


// ------------------------------------------------------------------------
// Local expansion, projection half: coef[l][t] += sum_j B_l(x_j) c_j,t with
// x_j = r_j / r_end.  Integer orders follow the reference's Wimp expansion
// (src/expansions.c:242-273): even nu -> eps_l T_2l(x), odd nu -> 2 T_2l+1(x).
// Half-integer orders (Chebyshev interpolation of j_n, not the reference's form) use
// T_2l+p(x), p = n mod 2, with weights c_j sqrt(r_j).
//
// The contraction is a (K+1) x count x 16 matrix product per item, run on
// the FP64 matrix cores: A[l][q] = B_l(x_q) (NLT row tiles of 16), B[q][t] =
// c_q,t, accumulated in registers over every source of the item; a pass
// stages kLocSub sources (basis by a three-term recurrence in steps of two
// degrees, one source per lane; coefficients loaded column by column,
// coalesced) in LDS with odd row
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

// LR: the contraction's precision (double, or float in mixed precision: the
// basis is computed in FP64 and rounded, the products and their sums over
// the item's sources are FP32, the per-item partial sums are added to the
// FP64 moments).
template <int NLT, typename LR>
__global__ __launch_bounds__(256) void k_loc_proj(const TfDesc *__restrict__ tfs,
                                                  const LocBoxD *__restrict__ boxes,
                                                  const LocItem *__restrict__ items,
                                                  double *__restrict__ coef)
{
    constexpr int LT = LocProjLds<NLT>::LT, LC = kLocBM + 1;
    extern __shared__ double sm_raw[];
    LR *sm = (LR *)sm_raw;
    LR *sT = sm;                              // [kLocSub][LT]
    LR *sC = sm + (size_t)kLocSub * LT;       // [kLocSub][LC]
    LocItem it = items[blockIdx.x];
    LocBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    const int K = T.K_loc, par = T.order & 1;
    const int nb = min(kLocBM, T.B - it.col0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    typename Mfma<LR>::V d[NLT];
#pragma unroll
    for (int lt = 0; lt < NLT; ++lt)
        for (int v = 0; v < 4; ++v) d[lt][v] = (LR)0;
    for (int s0 = 0; s0 < it.count; s0 += kLocSub) {
        const int n = min(kLocSub, it.count - s0);
        __syncthreads();
        if (tid < kLocSub) {
            LR *row = sT + (size_t)tid * LT;
            if (tid < n) {
                double x = ld_global(&T.r[bx.j0 + it.start + s0 + tid]) / bx.r_end;
                x = fmin(1.0, fmax(-1.0, x));
                // a_l = T_{2l+par}(x) by a_{l+1} = 2 T_2 a_l - a_{l-1} (from
                // T_m T_n = (T_{m+n} + T_{|m-n|}) / 2), a_{-1} = T_{par-2} =
                // T_{2-par}: one step per entry
                const double ty = 2.0 * (2.0 * x * x - 1.0);
                double am1 = par ? x : 0.5 * ty, a0 = par ? x : 1.0;
                row[0] = (LR)(par ? 2.0 * x : 1.0);
                for (int l = 1; l < 16 * NLT; ++l) {
                    const double a1 = fma(ty, a0, -am1);
                    am1 = a0;
                    a0 = a1;
                    row[l] = l <= K ? (LR)(2.0 * a0) : (LR)0;
                }
            } else {
                for (int l = 0; l < 16 * NLT; ++l) row[l] = (LR)0;
            }
        }
        // coefficients: lanes run over sources, 256 / kLocSub columns at a time
        for (int e = tid; e < kLocSub * kLocBM; e += 256) {
            int q = e % kLocSub, t = e / kLocSub;
            double c = 0.0;
            if (q < n && t < nb) {
                int j = bx.j0 + it.start + s0 + q;
                c = ld_global(&T.cs[(long long)(it.col0 + t) * T.ld_cs + j]);
                if (T.half) c *= sqrt(ld_global(&T.r[j]));
            }
            sC[q * LC + t] = (LR)c;
        }
        __syncthreads();
        const int ia = lane & 15, ka = lane >> 4;
        for (int st = wave; 4 * st < n; st += 4) {
            const int q = 4 * st + ka;
            const LR bv = sC[q * LC + ia];
#pragma unroll
            for (int lt = 0; lt < NLT; ++lt)
                d[lt] = Mfma<LR>::mma(sT[q * LT + 16 * lt + ia], bv, d[lt]);
        }
    }
    // sum the 4 waves' tiles: red[wave][16 NLT][LC]
    __syncthreads();
    LR *red = sm;
#pragma unroll
    for (int lt = 0; lt < NLT; ++lt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int row = 16 * lt + Mfma<LR>::row(lane, v);
            red[((size_t)wave * 16 * NLT + row) * LC + (lane & 15)] = d[lt][v];
        }
    __syncthreads();
    for (int e = tid; e < 16 * NLT * 16; e += 256) {
        int row = e >> 4, t = e & 15;
        if (row > K || t >= nb) continue;
        LR acc = (LR)0;
#pragma unroll
        for (int w = 0; w < 4; ++w) acc += red[((size_t)w * 16 * NLT + row) * LC + t];
        atomicAdd(&coef[bx.coef_off + (long long)row * T.B + it.col0 + t], (double)acc);
    }
}

// Local expansion, evaluation half (reference src/expansions.c:275-336):
// integer orders compute J_0..J_{l0+K+odd}(z), z = w r_end / 2, by one
// Miller sweep and form b_l = J_{l0+odd+l}(z) J_{|l0-l|}(z) (sign (-1) for
// l > l0 with l - l0 odd), then g[i,t] += sum_l b_l coef[l][t] for every
// column t; half-integer orders sample j_n(w r_end x_q) at the K+1 positive
// Chebyshev nodes, whose map to coefficients (the precomputed matrix
// cheb[]) is folded into the box's moments beforehand (k_loc_fold), and
// contract the samples with them.  One thread per target.  k_loc_eval1: one
// local box per item; the sums go out four (integer) or eight (half-integer)
// columns at a time.  k_loc_eval: every box that covers the item's targets
// (about 2.6 at the clustered workload), looped over with up to BMAX column
// sums in registers, so that each output element gets one atomic add per
// item instead of one per box (the registers cost occupancy, so transforms
// whose targets mostly meet one box keep the first form).  The per-target
// scratch lives in LDS as [slot][thread] (slots = loc_lds_slots(); none in
// k_loc_eval1's half-integer path), so lanes touch consecutive words.  A
// box's coefficients are the same for the whole block and are read as
// uniform loads.
__host__ __device__ inline int loc_slots(int order, int half, int K)
{
    return half ? 2 * (K + 1) : (order >> 1) + K + (order & 1) + 1;
}
// LDS words per target the evaluation kernels use: for half-integer orders
// the K + 1 node values (multi-box items; k_loc_eval1 keeps them in registers)
__host__ __device__ inline int loc_lds_slots(int order, int half, int K)
{
    return half ? K + 1 : loc_slots(order, half, K);
}

// j_n(hi + lo) out of line for k_loc_eval1's half-integer path (once per
// Chebyshev node): inlined there it raised the kernel's register count, and
// with it the integer orders' occupancy
__device__ __noinline__ double sph_jn_node(int n, double hi, double lo)
{
    return sph_jn_cw(n, hi, lo);
}

template <typename LR>
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
    const double sc0 = T.sign;
    const int B = T.B;
    if (!T.half) {
        // J_0..J_maxo by one Miller sweep, left unnormalised in S (the
        // normalisation 1/(sum + j) is applied to both factors of each b_l
        // in the contraction: no separate LDS pass)
        const int l0 = T.order >> 1, odd = T.order & 1;
        const int maxo = l0 + K + odd;
        const double z = w * bx.r_end * 0.5;
        double inv = 1.0;
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
            inv = 1.0 / (sum + j);
        }
        // b_l = J_{l0+odd+l} J_{|l0-l|} (sign (-1) for l > l0 with l - l0 odd)
        for (int t0 = 0; t0 < B; t0 += 4) {
            const int nt = min(4, B - t0);
            LR a0 = (LR)0, a1 = (LR)0, a2 = (LR)0, a3 = (LR)0;
            const double *cf = coef + bx.coef_off + t0;
            for (int l = 0; l <= K; ++l) {
                int lo = l0 - l;
                lo = lo < 0 ? -lo : lo;
                double v = (S[(l0 + odd + l) * nthr] * inv) * (S[lo * nthr] * inv);
                if (l > l0 && ((l - l0) & 1)) v = -v;
                const LR bl = (LR)v;
                const double *cr = cf + (long long)l * B;
                a0 = fma(bl, (LR)cr[0], a0);
                if (nt > 1) a1 = fma(bl, (LR)cr[1], a1);
                if (nt > 2) a2 = fma(bl, (LR)cr[2], a2);
                if (nt > 3) a3 = fma(bl, (LR)cr[3], a3);
            }
            double *g = T.gs + (long long)t0 * T.ld_gs + i;
            atomic_add_global(g, sc0 * (double)a0);
            if (nt > 1) atomic_add_global(g + T.ld_gs, sc0 * (double)a1);
            if (nt > 2) atomic_add_global(g + 2 * (long long)T.ld_gs, sc0 * (double)a2);
            if (nt > 3) atomic_add_global(g + 3 * (long long)T.ld_gs, sc0 * (double)a3);
        }
        return;
    }
    // half-integer orders: g_t = sign sqrt(2 w / pi) sum_q j_n(z x_q) C'[q][t]
    // at the nodes x_q = cos(pi (q + 1/2) / (2K + 2)), q = 0..K, with the
    // moments folded with the interpolation matrix (k_loc_fold: C' = M^T
    // coef, so that the sum equals sum_l b_l coef[l][t] with b_l = sum_q
    // M[l][q] j_n(z x_q)); no scratch, 8 columns per pass over the nodes
    const double z = w * bx.r_end;
    const double *xn = cheb + T.cheb_off + (K + 1) * (K + 1);   // the nodes, after the matrix
    const double sc = sc0 * sqrt(2.0 * w / kPi);
    const double *cf = coef + bx.coef_off;
    for (int t0 = 0; t0 < B; t0 += 8) {
        const int nt = min(8, B - t0);
        LR a[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) a[t] = (LR)0;
        for (int q = 0; q <= K; ++q) {
            const double xq = xn[q];
            const double hi = z * xq, lo = fma(z, xq, -hi);
            const LR f = (LR)sph_jn_node(T.order, hi, lo);
            const double *cr = cf + (long long)q * B + t0;
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t < nt) a[t] = fma(f, (LR)cr[t], a[t]);
        }
        double *g = T.gs + (long long)t0 * T.ld_gs + i;
#pragma unroll
        for (int t = 0; t < 8; ++t)
            if (t < nt) atomic_add_global(g + (long long)t * T.ld_gs, sc * (double)a[t]);
    }
}

// Half-integer orders: a local box's moments folded with the Chebyshev
// interpolation matrix of its transform, in place,
//   C'[q][t] = sum_l M[l][q] coef[l][t],  q, l = 0..K,
// so that the evaluation contracts the node values j_n(z x_q) directly.
// One block per box, kFoldCols columns at a time through LDS.
constexpr int kFoldCols = 32;
__global__ __launch_bounds__(256) void k_loc_fold(const TfDesc *__restrict__ tfs,
                                                  const LocBoxD *__restrict__ boxes,
                                                  const int *__restrict__ list,
                                                  double *__restrict__ coef,
                                                  const double *__restrict__ cheb)
{
    __shared__ double s_c[16 * kLocMaxNlt * kFoldCols];
    const LocBoxD bx = boxes[list[blockIdx.x]];
    const TfDesc &T = tfs[bx.tf];
    const int n = T.K_loc + 1, B = T.B;
    const double *M = cheb + T.cheb_off;
    double *c = coef + bx.coef_off;
    for (int t0 = 0; t0 < B; t0 += kFoldCols) {
        const int nc = min(kFoldCols, B - t0);
        __syncthreads();
        for (int o = threadIdx.x; o < n * nc; o += 256) {
            const int l = o / nc, t = o - l * nc;
            s_c[l * kFoldCols + t] = c[(long long)l * B + t0 + t];
        }
        __syncthreads();
        for (int o = threadIdx.x; o < n * nc; o += 256) {
            const int q = o / nc, t = o - q * nc;
            double a = 0.0;
            for (int l = 0; l < n; ++l) a = fma(M[(long long)l * n + q], s_c[l * kFoldCols + t], a);
            c[(long long)q * B + t0 + t] = a;
        }
    }
}

template <int BMAX, typename LR>
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
    LR acc[BMAX];
#pragma unroll
    for (int t = 0; t < BMAX; ++t) acc[t] = (LR)0;
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
            // j_n(z x_q) at the nodes, slots q, contracted below with the
            // folded moments C'[q][t] (k_loc_fold) as b_l with coef[l][t]
            const double z = w * bx.r_end;
            const double *xn = cheb + T.cheb_off + (K + 1) * (K + 1);
            for (int qq = 0; qq <= K; ++qq) {
                const double xq = xn[qq];
                const double hi = z * xq, lo = fma(z, xq, -hi);
                S[qq * nthr] = sph_jn_cw(T.order, hi, lo);
            }
            boff = 0;
        }
        const double *cf = coef + bx.coef_off + it.col0;
        for (int l = 0; l <= K; ++l) {
            const LR bl = (LR)S[(boff + l) * nthr];
            const double *cr = cf + (long long)l * B;
#pragma unroll
            for (int t = 0; t < BMAX; ++t)
                if (t < nb) acc[t] = fma(bl, (LR)cr[t], acc[t]);
        }
    }
    const double sc = T.sign * (T.half ? sqrt(2.0 * w / kPi) : 1.0);
    double *g = T.gs + (long long)it.col0 * T.ld_gs + i;
#pragma unroll
    for (int t = 0; t < BMAX; ++t)
        if (t < nb) atomic_add_global(g + (long long)t * T.ld_gs, sc * (double)acc[t]);
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
