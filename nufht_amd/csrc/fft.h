// Large batched complex FFTs of the asymptotic stage's grids, written for
// gfx950: the backward transform (e^{+2 pi i k n / Nf}, unnormalised) of
// every channel of an FFT size class, Nf = N1 * N2, in two passes over HBM
// (the four-step decomposition):
//
//   pass 1 (in place):     for each column n2, X1[k1][n2] = w^{k1 n2} *
//                          sum_n1 x[n1 N2 + n2] e^{2 pi i k1 n1 / N1}
//                          with w = e^{2 pi i / Nf}, stored at k1 N2 + n2;
//   pass 2 (out of place): for each row k1, X[k1 + N1 k2] =
//                          sum_n2 X1[k1][n2] e^{2 pi i k2 n2 / N2}.
//
// A block holds M whole sequences (M adjacent columns, or M adjacent rows)
// in LDS, so that global reads and writes move runs of M contiguous
// elements; within LDS each sequence is transformed by a Stockham radix-
// 2/3/4/8 FFT whose butterflies are held in registers across one barrier
// per stage (a single LDS buffer).
#pragma once
#include "device_common.h"
#include "fft_plan.h"

namespace nufht_dev {

template <typename R> struct CMul {
    typedef typename Cplx<R>::T C2;
    static __device__ __forceinline__ C2 mul(C2 a, C2 b)
    {
        C2 z;
        z.x = a.x * b.x - a.y * b.y;
        z.y = a.x * b.y + a.y * b.x;
        return z;
    }
    static __device__ __forceinline__ C2 add(C2 a, C2 b) { a.x += b.x; a.y += b.y; return a; }
    static __device__ __forceinline__ C2 sub(C2 a, C2 b) { a.x -= b.x; a.y -= b.y; return a; }
    static __device__ __forceinline__ C2 mul_i(C2 a) { C2 z; z.x = -a.y; z.y = a.x; return z; }
};

// Twiddles w^r, r = 1 .. RAD - 1, of a butterfly into x[1 .. RAD - 1]
// (the table holds w^m at index m).  FP64: one table load, the powers by
// complex products (w^(2s) = (w^s)^2, w^(2s+1) = w^(2s) w; at most four
// roundings deep, about 1e-15), which costs less than RAD - 1 loads through
// the vector cache.  FP32 keeps one load per power: its products' rounding
// would count against the mixed mode's error budget.
template <typename R, int RAD>
__device__ __forceinline__ void fft_twiddle(typename Cplx<R>::T (&x)[RAD],
                                            const typename Cplx<R>::T *__restrict__ W, int m)
{
    typedef typename Cplx<R>::T C2;
    if constexpr (sizeof(R) == 8) {
        C2 w[RAD];
        w[1] = W[m];
#pragma unroll
        for (int r = 2; r < RAD; ++r) w[r] = CMul<R>::mul(w[r >> 1], w[(r + 1) >> 1]);
#pragma unroll
        for (int r = 1; r < RAD; ++r) x[r] = CMul<R>::mul(x[r], w[r]);
    } else {
#pragma unroll
        for (int r = 1; r < RAD; ++r) x[r] = CMul<R>::mul(x[r], W[r * m]);
    }
}

// small DFTs with the backward sign
template <typename R, int RAD> struct Dft;
template <typename R> struct Dft<R, 2> {
    typedef typename Cplx<R>::T C2;
    static __device__ __forceinline__ void run(C2 *x)
    {
        C2 a = x[0], b = x[1];
        x[0] = CMul<R>::add(a, b);
        x[1] = CMul<R>::sub(a, b);
    }
};
template <typename R> struct Dft<R, 3> {
    typedef typename Cplx<R>::T C2;
    static __device__ __forceinline__ void run(C2 *x)
    {
        const R h = (R)0.86602540378443864676;   // sqrt(3) / 2
        C2 t = CMul<R>::add(x[1], x[2]), d = CMul<R>::sub(x[1], x[2]);
        C2 m;
        m.x = x[0].x - (R)0.5 * t.x;
        m.y = x[0].y - (R)0.5 * t.y;
        x[0] = CMul<R>::add(x[0], t);
        C2 id;                                    // i (sqrt(3)/2) (x1 - x2)
        id.x = -h * d.y;
        id.y = h * d.x;
        x[1] = CMul<R>::add(m, id);
        x[2] = CMul<R>::sub(m, id);
    }
};
template <typename R> struct Dft<R, 4> {
    typedef typename Cplx<R>::T C2;
    static __device__ __forceinline__ void run(C2 *x)
    {
        C2 a = CMul<R>::add(x[0], x[2]), b = CMul<R>::sub(x[0], x[2]);
        C2 c = CMul<R>::add(x[1], x[3]), d = CMul<R>::mul_i(CMul<R>::sub(x[1], x[3]));
        x[0] = CMul<R>::add(a, c);
        x[2] = CMul<R>::sub(a, c);
        x[1] = CMul<R>::add(b, d);
        x[3] = CMul<R>::sub(b, d);
    }
};
// 9 = 3 x 3: x[3 n1 + n2] -> 3-point DFTs over n1, twiddles e^{2 pi i n2 k1 / 9},
// 3-point DFTs over n2 into X[k1 + 3 k2]
template <typename R> struct Dft<R, 9> {
    typedef typename Cplx<R>::T C2;
    static __device__ __forceinline__ void run(C2 *x)
    {
        C2 y[3][3];
#pragma unroll
        for (int n2 = 0; n2 < 3; ++n2) {
            C2 t[3] = {x[n2], x[3 + n2], x[6 + n2]};
            Dft<R, 3>::run(t);
            y[n2][0] = t[0];
            y[n2][1] = t[1];
            y[n2][2] = t[2];
        }
        C2 w1, w2, w4;                            // e^{2 pi i m / 9}, m = 1, 2, 4
        w1.x = (R)0.76604444311897803520;         // cos(2 pi / 9)
        w1.y = (R)0.64278760968653932632;         // sin(2 pi / 9)
        w2.x = (R)0.17364817766693034885;         // cos(4 pi / 9)
        w2.y = (R)0.98480775301220805936;         // sin(4 pi / 9)
        w4.x = (R)-0.93969262078590838405;        // cos(8 pi / 9)
        w4.y = (R)0.34202014332566873304;         // sin(8 pi / 9)
        y[1][1] = CMul<R>::mul(y[1][1], w1);
        y[1][2] = CMul<R>::mul(y[1][2], w2);
        y[2][1] = CMul<R>::mul(y[2][1], w2);
        y[2][2] = CMul<R>::mul(y[2][2], w4);
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
            C2 t[3] = {y[0][k1], y[1][k1], y[2][k1]};
            Dft<R, 3>::run(t);
            x[k1] = t[0];
            x[k1 + 3] = t[1];
            x[k1 + 6] = t[2];
        }
    }
};
template <typename R> struct Dft<R, 8> {
    typedef typename Cplx<R>::T C2;
    static __device__ __forceinline__ void run(C2 *x)
    {
        const R s = (R)0.70710678118654752440;   // sqrt(2) / 2
        C2 e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
        Dft<R, 4>::run(e);
        Dft<R, 4>::run(o);
        // o[k] *= e^{i pi k / 4}
        C2 t;
        t.x = s * (o[1].x - o[1].y);
        t.y = s * (o[1].x + o[1].y);
        o[1] = t;
        o[2] = CMul<R>::mul_i(o[2]);
        t.x = -s * (o[3].x + o[3].y);
        t.y = s * (o[3].x - o[3].y);
        o[3] = t;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k] = CMul<R>::add(e[k], o[k]);
            x[k + 4] = CMul<R>::sub(e[k], o[k]);
        }
    }
};

// LDS position of element n of sequence c: rows of stride = M + 1 and one
// more pad element every 2^shift = 8 rows, so that consecutive n (the row
// pass's loads) and the stride-R writes of a first stage both spread over the
// banks; where that padding would keep a second block off the CU (more than
// 80 KiB), rows of stride = M and one pad element every 2^shift = 4 rows.
__device__ __forceinline__ int fft_lds_idx(int n, int c, int stride, int shift)
{
    return n * stride + c + (n >> shift);
}

// One Stockham stage of radix RAD over M sequences of length N held in LDS
// at fft_lds_idx(n, c); Ns = product of the earlier radices.
// Ns is a power of 2 (lgNs its log): the radix-3 or radix-9 stage, if any,
// comes last.
template <typename R, int RAD>
__device__ __forceinline__ void fft_stage(typename Cplx<R>::T *s, int N, int M, int lgM, int lgNs,
                                          int ls, int lh, const typename Cplx<R>::T *__restrict__ W)
{
    typedef typename Cplx<R>::T C2;
    constexpr int MAXB = (fft_max_elems<R>() / RAD + kFftThreads - 1) / kFftThreads;
    const int Ns = 1 << lgNs;
    const int nb = M * N / RAD, span = N / RAD;
    const int wstep = N / (Ns * RAD);
    C2 x[MAXB][RAD];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        int b = threadIdx.x + q * kFftThreads;
        if (b < nb) {
            int c = b & (M - 1), j = b >> lgM, k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < RAD; ++r) x[q][r] = s[fft_lds_idx(j + r * span, c, ls, lh)];
            if (Ns > 1) fft_twiddle<R, RAD>(x[q], W, k * wstep);
            Dft<R, RAD>::run(x[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        int b = threadIdx.x + q * kFftThreads;
        if (b < nb) {
            int c = b & (M - 1), j = b >> lgM, k = j & (Ns - 1);
            int base = (j >> lgNs) * Ns * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) s[fft_lds_idx(base + r * Ns, c, ls, lh)] = x[q][r];
        }
    }
    __syncthreads();
}

template <typename R>
__device__ __forceinline__ void fft_lds(typename Cplx<R>::T *s, const FftPass &P,
                                        const typename Cplx<R>::T *__restrict__ W)
{
    int lgNs = 0;
    for (int st = 0; st < P.nstage; ++st) {
        int rad = P.radix[st];
        if (rad == 8) fft_stage<R, 8>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else if (rad == 4) fft_stage<R, 4>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else if (rad == 3) fft_stage<R, 3>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else if (rad == 9) fft_stage<R, 9>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else fft_stage<R, 2>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        lgNs += rad == 8 ? 3 : rad == 4 ? 2 : 1;
    }
}

// The first Stockham stage (Ns = 1) with its inputs taken straight from
// global memory (ld(c, n): element n of sequence c) and its outputs written
// to LDS, and the last stage (Ns = N / RAD) with its outputs handed to st(c,
// n, v) instead of LDS: two LDS passes and one barrier fewer per block.
// All of a thread's loads are issued before the first butterfly.
template <typename R, int RAD, class Load>
__device__ __forceinline__ void fft_first_stage(typename Cplx<R>::T *s, const FftPass &P, Load ld)
{
    typedef typename Cplx<R>::T C2;
    constexpr int MAXB = (fft_max_elems<R>() / RAD + kFftThreads - 1) / kFftThreads;
    const int nb = P.M * P.N / RAD, span = P.N / RAD;
    C2 x[MAXB][RAD];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * kFftThreads;
        if (b < nb) {
            const int c = b & (P.M - 1), j = b >> P.lgM;
#pragma unroll
            for (int r = 0; r < RAD; ++r) x[q][r] = ld(c, j + r * span);
        }
    }
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * kFftThreads;
        if (b < nb) {
            const int c = b & (P.M - 1), j = b >> P.lgM;
            Dft<R, RAD>::run(x[q]);
#pragma unroll
            for (int r = 0; r < RAD; ++r) s[fft_lds_idx(j * RAD + r, c, P.lds_stride, P.lds_shift)] = x[q][r];
        }
    }
    __syncthreads();
}

template <typename R, int RAD, class Store>
__device__ __forceinline__ void fft_last_stage(const typename Cplx<R>::T *s, const FftPass &P,
                                               const typename Cplx<R>::T *__restrict__ W, Store st)
{
    typedef typename Cplx<R>::T C2;
    constexpr int MAXB = (fft_max_elems<R>() / RAD + kFftThreads - 1) / kFftThreads;
    const int Ns = P.N / RAD, nb = P.M * Ns;
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * kFftThreads;
        if (b < nb) {
            const int c = b & (P.M - 1), j = b >> P.lgM;     // j < Ns: k = j
            C2 x[RAD];
#pragma unroll
            for (int r = 0; r < RAD; ++r) x[r] = s[fft_lds_idx(j + r * Ns, c, P.lds_stride, P.lds_shift)];
            fft_twiddle<R, RAD>(x, W, j);
            Dft<R, RAD>::run(x);
#pragma unroll
            for (int r = 0; r < RAD; ++r) st(c, j + r * Ns, x[r]);
        }
    }
}

// The whole transform of a tile: first stage from global memory, the middle
// stages in LDS, the last stage to global memory (at least two stages).
template <typename R, class Load, class Store>
__device__ __forceinline__ void fft_tile(typename Cplx<R>::T *s, const FftPass &P,
                                         const typename Cplx<R>::T *__restrict__ W, Load ld, Store st)
{
    const int r0 = P.radix[0];
    if (r0 == 8) fft_first_stage<R, 8>(s, P, ld);
    else if (r0 == 4) fft_first_stage<R, 4>(s, P, ld);
    else fft_first_stage<R, 2>(s, P, ld);
    int lgNs = r0 == 8 ? 3 : r0 == 4 ? 2 : 1;
    for (int q = 1; q + 1 < P.nstage; ++q) {
        const int rad = P.radix[q];
        if (rad == 8) fft_stage<R, 8>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else if (rad == 4) fft_stage<R, 4>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else if (rad == 3) fft_stage<R, 3>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else if (rad == 9) fft_stage<R, 9>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        else fft_stage<R, 2>(s, P.N, P.M, P.lgM, lgNs, P.lds_stride, P.lds_shift, W);
        lgNs += rad == 8 ? 3 : rad == 4 ? 2 : 1;
    }
    const int rl = P.radix[P.nstage - 1];
    if (rl == 8) fft_last_stage<R, 8>(s, P, W, st);
    else if (rl == 4) fft_last_stage<R, 4>(s, P, W, st);
    else if (rl == 3) fft_last_stage<R, 3>(s, P, W, st);
    else if (rl == 9) fft_last_stage<R, 9>(s, P, W, st);
    else fft_last_stage<R, 2>(s, P, W, st);
}

// Blocks go to the 8 XCDs (each with its own L2) round robin in launch
// order; a tile's runs of M elements fill half of a 128-byte line when M
// elements are 64 bytes, the other half belonging to the next tile.  Within
// each group of 16 blocks, blocks b and b + 8 (same XCD) take adjacent tiles,
// so that a line's two halves meet in one L2.
__device__ __forceinline__ int fft_block_tile()
{
    const int b = blockIdx.x;
    if (b >= (int)(gridDim.x & ~15u)) return b;
    return (b & ~15) + 2 * (b & 7) + ((b >> 3) & 1);
}

// Pass 1: column FFTs of length N1 (stride N2) in place, times w^{k1 n2}.
template <typename R>
__global__ __launch_bounds__(kFftThreads, 4) void k_fft_cols(typename Cplx<R>::T *__restrict__ g,
                                                          FftPass P,
                                                          const typename Cplx<R>::T *__restrict__ tw,
                                                          const int4 *__restrict__ chinfo)
{
    typedef typename Cplx<R>::T C2;
    extern __shared__ unsigned char s_raw[];
    C2 *s = (C2 *)s_raw;
    const int bt = fft_block_tile();
    const int ch = bt / P.tiles, t = bt % P.tiles;
    const int M = P.M;
    const int L = chinfo[P.chinfo_off + ch].x;
    const int nf = (int)P.nf;
    C2 *base = g + (long long)ch * P.nf + (long long)t * M;
    const C2 *hi = tw + P.hi_off, *lo = tw + P.lo_off;
    // element n of column c: grid position p = n N2 + t M + c, zero (not
    // stored) outside |l| <= L; out: times w^{k1 (t M + c)}, in place
    auto ld = [&](int c, int n) {
        const int p = n * P.N2 + t * M + c;
        C2 v;
        v.x = (R)0;
        v.y = (R)0;
        if (p <= L || p >= nf - L) v = base[(long long)n * P.N2 + c];
        return v;
    };
    // The last stage hands a thread's outputs k1 = j + r Ns, r = 0, 1, ...
    // (Ns = N1 / its radix) in that order: in FP64 the twiddles w^{k1 n2}
    // follow from w^{j n2} and w^{Ns n2} (two table products) by a running
    // product, instead of two table loads per element; FP32 grids read the
    // table for every element.
    const int Ns = P.N / P.radix[P.nstage - 1];
    C2 twc, tws;
    auto st = [&](int c, int k1, C2 v) {
        const int n2 = t * M + c;
        if constexpr (sizeof(R) == 8) {
            if (k1 < Ns) {
                const int a = k1 * n2, b = Ns * n2;   // < Ns N2 <= Nf
                twc = CMul<R>::mul(hi[a >> 8], lo[a & 255]);
                tws = CMul<R>::mul(hi[b >> 8], lo[b & 255]);
            } else {
                twc = CMul<R>::mul(twc, tws);
            }
        } else {
            const int a = k1 * n2;                  // < (N1 - 1) (N2 - 1) < Nf
            twc = CMul<R>::mul(hi[a >> 8], lo[a & 255]);
        }
        base[(long long)k1 * P.N2 + c] = CMul<R>::mul(v, twc);
    };
    fft_tile<R>(s, P, tw + P.w_off, ld, st);
}

// Pass 2: row FFTs of length N2, written to natural order in out.
template <typename R>
__global__ __launch_bounds__(kFftThreads, 4) void k_fft_rows(const typename Cplx<R>::T *__restrict__ g,
                                                          typename Cplx<R>::T *__restrict__ out,
                                                          FftPass P,
                                                          const typename Cplx<R>::T *__restrict__ tw,
                                                          const int4 *__restrict__ chinfo)
{
    typedef typename Cplx<R>::T C2;
    extern __shared__ unsigned char s_raw[];
    C2 *s = (C2 *)s_raw;
    const int bt = fft_block_tile();
    const int ch = bt / P.tiles, t = bt % P.tiles;
    const int M = P.M;
    const C2 *src = g + (long long)ch * P.nf + (long long)t * M * P.N2;
    C2 *dst = out + (long long)ch * P.nf + (long long)t * M;
    const int4 ci = chinfo[P.chinfo_off + ch];
    const int nf = (int)P.nf;
    auto ld = [&](int c, int n) { return src[(long long)c * P.N2 + n]; };
    // output k2 of row t M + c is frequency q = t M + c + k2 N1; only the
    // interpolation window [mlo, mlo + span] (mod Nf) is written
    auto st = [&](int c, int k2, C2 v) {
        const int q = t * M + c + k2 * P.N1;
        int d = q - ci.y;                  // mlo in (-Nf, Nf)
        d = d < 0 ? d + nf : (d >= nf ? d - nf : d);
        if (d <= ci.z) dst[(long long)k2 * P.N1 + c] = v;
    };
    fft_tile<R>(s, P, tw + P.w_off, ld, st);
}

// Lengths up to one tile (Nf <= 4096 in FP64, 8192 in FP32): one pass, M
// whole channels per block (each a contiguous row), the same Stockham stages
// in LDS.  Grid positions outside [-L, L] (mod Nf) are zeros the spread did
// not store; only the frequencies mlo .. mlo + span (mod Nf) are written.
template <typename R>
__global__ __launch_bounds__(kFftThreads) void k_fft_small(const typename Cplx<R>::T *__restrict__ g,
                                                           typename Cplx<R>::T *__restrict__ out,
                                                           FftPass P,
                                                           const typename Cplx<R>::T *__restrict__ tw,
                                                           const int4 *__restrict__ chinfo)
{
    typedef typename Cplx<R>::T C2;
    extern __shared__ unsigned char s_raw[];
    C2 *s = (C2 *)s_raw;
    const int N = P.N, M = P.M, ch0 = blockIdx.x * M;
    for (int idx = threadIdx.x; idx < M * N; idx += kFftThreads) {
        const int c = idx / N, n = idx - c * N, ch = ch0 + c;
        C2 v;
        v.x = (R)0;
        v.y = (R)0;
        if (ch < P.nch) {
            const int L = chinfo[P.chinfo_off + ch].x;
            if (n <= L || n >= N - L) v = g[(long long)ch * N + n];
        }
        s[fft_lds_idx(n, c, P.lds_stride, P.lds_shift)] = v;
    }
    __syncthreads();
    fft_lds<R>(s, P, tw + P.w_off);
    for (int idx = threadIdx.x; idx < M * N; idx += kFftThreads) {
        const int c = idx / N, n = idx - c * N, ch = ch0 + c;
        if (ch >= P.nch) continue;
        const int4 ci = chinfo[P.chinfo_off + ch];
        int d = n - ci.y;
        d = d < 0 ? d + N : (d >= N ? d - N : d);
        if (d <= ci.z) out[(long long)ch * N + n] = s[fft_lds_idx(n, c, P.lds_stride, P.lds_shift)];
    }
}

}  // namespace nufht_dev
