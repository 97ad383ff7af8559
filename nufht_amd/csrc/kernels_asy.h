// Asymptotic-stage kernels: grid zeroing, spreading (VALU and matrix-core)
// and interpolation (VALU and matrix-core).  Instantiated per (grid
// precision, exponent bound) in asy_stage.hip.
#pragma once
#include "device_common.h"

namespace nufht_dev {


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
// tile.  A lane owns one position, two columns and EG consecutive exponents
// (spread_valu_lanes gives the lanes per position; the host sizes the tile
// so that they fit in 256): the lanes of an exponent group and column pair
// take consecutive positions, so that every store writes runs of adjacent
// grid points, and a lane holds 4 EG accumulators whatever the box's E.
template <typename R, int EMAX>
__global__ __launch_bounds__(256, 4) void k_asy_spread(const TfDesc *__restrict__ tfs,
                                                    const AsyBoxD *__restrict__ boxes,
                                                    const AsyItem *__restrict__ items,
                                                    const int2 *__restrict__ ranges,
                                                    typename Cplx<R>::T *__restrict__ grid,
                                                    const double *__restrict__ ft)
{
    typedef typename Cplx<R>::T C2;
    constexpr int EG = SpreadEG<R>::v;
    constexpr int NG = (EMAX + EG - 1) / EG;
    __shared__ C2 s_z[kVQ][kTapMax];       // psi * e^{iDx} r^{-1/2}, tap k of source q
    __shared__ C2 s_c[8][kVQ];             // column pair, source
    __shared__ R s_pw[kVQ][NG + 1];        // (rho/r)^(g EG) for group g; slot NG: rho/r
    __shared__ double s_x[kVQ];
    __shared__ double2 s_ph[kVQ];          // e^{iDx} r^{-1/2}
    __shared__ int s_pf[kVQ];
    __shared__ double s_dec[256];
    const AsyItem it = items[blockIdx.x];
    const int2 jr = ranges[blockIdx.x];
    const AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    const int tb = it.tb, tl2 = tb > 1 ? tb >> 1 : 1;
    const int E = bx.E, ng = (E + EG - 1) / EG;
    const int np = it.count;
    const int tid = threadIdx.x;
    const int pl = tid % np, combo = tid / np;
    const int tl = combo % tl2, g = combo / tl2;
    const bool lane_on = g < ng;
    const int p = it.start + pl;
    const int l = p <= bx.L ? p : p - bx.Nf;
    const bool valid = (l >= -bx.L) && (l <= bx.L);
    const int t0 = it.col0 + 2 * tl;
    const bool active = lane_on && valid && (t0 < T.B);
    const int kw = T.kw;
    const double inv_a1 = 1.0 / bx.a1;
    R a0r[EG], a0i[EG], a1r[EG], a1i[EG];
#pragma unroll
    for (int e = 0; e < EG; ++e) { a0r[e] = a0i[e] = a1r[e] = a1i[e] = (R)0; }
    for (int c0 = jr.x; c0 < jr.y; c0 += kVQ) {
        const int cn = min(kVQ, jr.y - c0);
        if (tid < cn) {
            const double rj = ld_global(&T.r[c0 + tid]);
            const double x = rj - bx.C;
            s_x[tid] = x;
            s_pf[tid] = (int)ceil((x - bx.a1) / bx.h);
            double sn, cs;
            sincos_cw<SC_CALL>(bx.D * x, &sn, &cs);
            const double rs = 1.0 / sqrt(rj);
            s_ph[tid] = make_double2(cs * rs, sn * rs);
            const double ri = bx.rho / rj;
            double rg = 1.0;
#pragma unroll
            for (int e = 0; e < EG; ++e) rg *= ri;
            double pw = 1.0;
#pragma unroll
            for (int gg = 0; gg < NG; ++gg) { s_pw[tid][gg] = (R)pw; pw *= rg; }
            s_pw[tid][NG] = (R)ri;
        }
        // coefficients: consecutive lanes load consecutive sources of a column
        for (int q = tid; q < tl2 * cn; q += 256) {
            const int pr = q / cn, jq = q - pr * cn;
            const int ta = it.col0 + 2 * pr, tb2 = ta + 1;
            C2 v;
            v.x = ta < T.B && 2 * pr < tb ? (R)ld_global(&T.cs[(long long)ta * T.ld_cs + c0 + jq]) : (R)0;
            v.y = tb2 < T.B && 2 * pr + 1 < tb ? (R)ld_global(&T.cs[(long long)tb2 * T.ld_cs + c0 + jq]) : (R)0;
            s_c[pr][jq] = v;
        }
        __syncthreads();
        for (int q = tid; q < cn * (kw + 1); q += 256) {
            const int jq = q / (kw + 1), k = q - jq * (kw + 1);
            const double x = s_x[jq];
            const double u = ((double)(s_pf[jq] + k) * bx.h - x) * inv_a1;
            const double psi = (double)es_eval((R)u, (R)T.beta);
            const double2 ph = s_ph[jq];
            C2 z;
            z.x = (R)(ph.x * psi);
            z.y = (R)(ph.y * psi);
            s_z[jq][k] = z;
        }
        __syncthreads();
        if (active) {
            // sources of this chunk whose taps reach l: lf in [l - kw, l] (lf ascending)
            int a = 0, b = cn;
            while (a < b) { int mid = (a + b) >> 1; if (s_pf[mid] >= l - kw) b = mid; else a = mid + 1; }
            const int qa = a;
            b = cn;
            while (a < b) { int mid = (a + b) >> 1; if (s_pf[mid] > l) b = mid; else a = mid + 1; }
            const int qb = a;
            for (int q = qa; q < qb; ++q) {
                const C2 z = s_z[q][l - s_pf[q]];
                const C2 cc = s_c[tl][q];
                const R w = s_pw[q][g], ri = s_pw[q][NG];
                const R zr = z.x * w, zi = z.y * w;
                const R z0r = zr * cc.x, z0i = zi * cc.x, z1r = zr * cc.y, z1i = zi * cc.y;
                R rp = (R)1;
#pragma unroll
                for (int e = 0; e < EG; ++e) {
                    a0r[e] = fma(z0r, rp, a0r[e]);
                    a0i[e] = fma(z0i, rp, a0i[e]);
                    a1r[e] = fma(z1r, rp, a1r[e]);
                    a1i[e] = fma(z1i, rp, a1i[e]);
                    rp *= ri;
                }
            }
        }
        __syncthreads();
    }
    // deconvolution once per position
    if (tid < np) {
        const int pq = it.start + tid;
        const int lq = pq <= bx.L ? pq : pq - bx.Nf;
        s_dec[tid] = (lq >= -bx.L && lq <= bx.L)
                         ? 1.0 / (bx.a2 * ft_eval(ft + T.ft_off, (double)lq * bx.h * bx.a2, T.kmax))
                         : 0.0;
    }
    __syncthreads();
    if (lane_on && t0 < T.B) {
        const double dec = s_dec[pl];
        const bool two = (t0 + 1 < T.B) && (2 * tl + 1 < tb);
#pragma unroll
        for (int e2 = 0; e2 < EG; ++e2) {
            const int e = g * EG + e2;
            if (e >= E) break;
            C2 *d0 = grid + bx.goff + ((long long)(e * T.B + t0)) * bx.Nf + p;
            const R v0r = (R)((double)a0r[e2] * dec), v0i = (R)((double)a0i[e2] * dec);
            const R v1r = (R)((double)a1r[e2] * dec), v1i = (R)((double)a1i[e2] * dec);
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
// Row strides are padded so that both the staging writes (one source per
// lane) and the product reads (4 sources x 4 positions or exponents per
// instruction) spread over the LDS banks.
constexpr int kSpreadDec = 128;   // deconvolution factors staged per chunk
template <typename R, int EMAX> struct SpreadLds {
    typedef typename Cplx<R>::T C2;
    R psi[kSQ][kSP + 5];          // kernel value of source q at position l: column l mod 16
    C2 u[kSQ][EMAX + 1];          // e^{iDx} r^{-1/2} (rho/r)^e
    R bop[kSQ][17];               // A operand: c[q][t] (rho/r)^{e2}, row j = (t, e2)
    double dec[kSpreadDec];       // 1 / (a2 F(l h a2)) for l = dec_base + i
    int lf[kSQ];                  // first kernel tap of source q
    int dec_base;
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

// Global loads of one producer thread for one chunk (its source's r; for
// waves 2-3 also 8 coefficients), issued a chunk ahead of their use.
template <bool CV> struct SpreadIn { static constexpr bool kCV = true; double r; double cv[8]; };
template <> struct SpreadIn<false> { static constexpr bool kCV = false; double r; };

template <bool CV>
__device__ __forceinline__ void spread_load(const TfDesc &T, const AsyItem &it, int c0, int cn,
                                            int ptid, SpreadIn<CV> &in)
{
    const int q = ptid & 63;
    in.r = ld_global(&T.r[c0 + min(q, cn - 1)]);
    if constexpr (CV) {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int j = (ptid - 128 + 128 * m) >> 6, tc = it.col0 + j % it.tb;
            in.cv[m] = (q < cn && tc < T.B) ? ld_global(&T.cs[(long long)tc * T.ld_cs + c0 + q]) : 0.0;
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
        if (it.pad) {
            // the deconvolution factors of positions dec_base .. dec_base +
            // 127, from 32 below the chunk's first tap: the window flushes
            // while the chunk is multiplied reach mostly those (others
            // evaluate their own)
            const double r0 = __shfl(in.r, 0, 64);
            const int base = (int)ceil((r0 - bx.C - bx.a1) / bx.h) - 32, i = ptid - 128;
            L.dec[i] = 1.0 / (bx.a2 * ft_eval(ft + T.ft_off, (double)(base + i) * bx.h * bx.a2, T.kmax));
            if (i == 0) L.dec_base = base;
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
        spread_stage<R, EMAX, CV>(T, bx, it, piece.x, min(kSQ, piece.y - piece.x), pw, ptid, in, ft,
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
            spread_stage<R, EMAX, CV>(T, bx, it, c1, min(kSQ, piece.y - c1), pw, ptid, in, ft,
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
    const SpreadLds<R, EMAX> *Lc = L;     // the chunk being multiplied (its staged factors)
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
            const int di = it.pad ? P - Lc->dec_base : -1;
            const double dec = di >= 0 && di < kSpreadDec
                                   ? Lc->dec[di]
                                   : 1.0 / (bx.a2 * ft_eval(ft + T.ft_off, (double)P * bx.h * bx.a2, T.kmax));
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
                spread_stage<R, EMAX, decltype(in)::kCV>(T, bx, it, c0, cn, wv, tid, in, ft, L[0]);
            } else {
                SpreadIn<false> in;
                spread_load(T, it, c0, cn, tid, in);
                spread_stage<R, EMAX, decltype(in)::kCV>(T, bx, it, c0, cn, wv, tid, in, ft, L[0]);
            }
            __syncthreads();
        }
        const SpreadLds<R, EMAX> &Lk = L[SPLIT ? (k & 1) : 0];
        Lc = &Lk;
        // the chunk's first taps in a register (lane q holds source q's),
        // read by lane index below instead of one LDS round trip each
        const int lfl = Lk.lf[min(lane, cn - 1)];
        auto lfq = [&](int q) { return __builtin_amdgcn_readlane(lfl, min(q, cn - 1)); };
        const int l0 = lfq(0);
        const int l1 = lfq(cn - 1);
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
            // (a step's operands are read before the window moves, so that
            // the reads overlap the flush)
            const int ka = lane >> 4, ja = lane & 15;
            for (int s = 0; s < cn; s += 4) {
                const int q = s + ka;
                const R a = Lk.bop[q][ja], pr = Lk.psi[q][ja];
                const R *uq = (const R *)&Lk.u[q][0];
                R uv[KW];
#pragma unroll
                for (int k = 0; k < KW; ++k) uv[k] = uq[uoff[k]];
                int lo = 0;
                while (lo < 4) {
                    const int llo = lfq(s + lo);
                    advance(llo + kw - 15);
                    int hi = lo + 1;
                    while (hi < 4 && lfq(s + hi) + kw <= W + 15) ++hi;
                    const R ps = (ka >= lo && ka < hi) ? pr : (R)0;
#pragma unroll
                    for (int k = 0; k < KW; ++k)
                        if (k < mc) acc[k] = Mfma<R>::mma(a, ps * uv[k], acc[k]);
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

// One launch per products-per-wave count KW (the host groups the items by
// spread_mfma_kw), so that each launch holds only the accumulators its items
// use: with KW <= 5 and the split schedule at most 128 VGPRs, so that two
// blocks share a CU (the few values that then leave registers are reloaded
// once per chunk).
template <typename R, int EMAX, bool SP, int KW>
__global__ __launch_bounds__(SpreadSched<SP>::threads, (SP && KW <= 5) ? 4 : 1) void k_asy_spread_mfma(
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
    const int nprod = 2 * ((bx.E + nsub - 1) / nsub);
    spread_mfma_body<R, EMAX, KW, SP>(T, bx, it, rg, nprod, grid, ft, L);
}


// Interpolation, target-dense runs, on the matrix cores.  An item is a run
// of up to kIrTiles consecutive 16-target tiles of one box whose kernel
// supports each fit a window of 4 kIrSteps grid positions, and whose windows all
// lie in [mw0, mw0 + npos), npos <= kIrPos.  Targets are dense against the
// grid (consecutive tiles mostly share their window: the headline and
// clustered workloads read each grid position 10 to 13 times as 16-target
// windows), so the run's positions are staged in LDS once per (column
// chunk, exponent group) with coalesced loads of whole channel rows, and
// every tile's products read their window from there.  For exponent group
// g (nsub = 16/tb exponents) and tile u:
//   D[i][(t, e2)] = sum_m phi(s_i - m Delta) grid[e][t][m],  e = g nsub + e2
// as 16x16x4 FP64 matrix products (up to kIrSteps K steps of 4 positions, the phi
// operand computed once per tile); the FP64 epilogue applies
// e^{i(phi0 + w C)} h Delta / psihat(s) and the Hankel weights.  The
// Hankel combination is folded into a running complex sum per (target,
// column): S = sum_e A'_e i^(e mod 2) G_e with A'_e = +-a_e (rho w)^-e, and
// the output is b0 Re(f S) with f = e^{i(phi0 + w C)} h Delta / psihat(s):
// the real part for even e and minus the imaginary part for odd e, as the
// reference's cos / sin terms (src/expansions.c:523-531, :549-557).
constexpr int kIrTargets = 16 * kIrTiles;
constexpr int kIrStride = kIrPos + 1;        // odd row stride: conflict-free reads

template <typename GR, typename CR, int EMAX>
__global__ __launch_bounds__(256) void k_asy_interp_mfma(const TfDesc *__restrict__ tfs,
                                                         const AsyBoxD *__restrict__ boxes,
                                                         const InterpRun *__restrict__ items,
                                                         const typename Cplx<GR>::T *__restrict__ grid,
                                                         const double *__restrict__ ft)
{
    typedef CR R;
    typedef typename Cplx<GR>::T C2;
    typedef typename Mfma<R>::V V;
    constexpr int TPW = kIrTiles / 4;
    __shared__ C2 s_g[16 * kIrStride];           // staged window: [channel][position]
    __shared__ double s_s[kIrTargets], s_fr[kIrTargets], s_fi[kIrTargets];
    __shared__ int s_m0[kIrTargets];
    __shared__ double s_P[kIrTargets][EMAX];
    __shared__ double s_out[4][16][17];           // per wave: result tile, transposed
    __shared__ __align__(16) double s_tap[kTapMax][16];   // kernel tap series, rows padded
    const InterpRun it = items[xcd_item<4>()];
    const AsyBoxD bx = boxes[it.box];
    const TfDesc &T = tfs[bx.tf];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int E = bx.E, kw = T.kw;
    for (int q = tid; q < kw * kEsN; q += 256) {
        const int k = q / kEsN;
        s_tap[k][q - k * kEsN] = ft[T.ft_off + kNFT + q];
    }
    // per-target factors, one thread per target (rows past the run's end
    // repeat its last target with zero weight); the Hankel weights' powers
    // on the block's other half
    static_assert(2 * kIrTargets == 256, "one thread per target in each half");
    if (tid >= kIrTargets) {
        const int ti = tid - kIrTargets;
        const double w = ld_global(&T.w[bx.i0 + it.start + min(ti, it.count - 1)]);
        const double rw = 1.0 / (bx.rho * w);
        double pw = 1.0;
        for (int e = 0; e < EMAX; ++e) {
            s_P[ti][e] = e < E ? ((((e >> 1) & 1) ? -T.a[e] : T.a[e]) * pw) : 0.0;
            pw *= rw;
        }
    } else {
        const int ti = tid;
        const double w = ld_global(&T.w[bx.i0 + it.start + min(ti, it.count - 1)]);
        const double s = w - bx.D;
        s_s[ti] = s;
        s_m0[ti] = (int)ceil((s - bx.a2) / bx.delta);
        const double nu = (double)T.order + 0.5 * (double)T.half;
        double sp, cp;
        sincos_cw(w * bx.C, &sp, &cp);
        const int q = (int)lround(2.0 * nu + 1.0) & 7;
        const double c0 = cos_qpi4(q), s0 = -cos_qpi4((q + 6) & 7);
        const double amp = bx.h * bx.delta / (bx.a1 * ft_eval(ft + T.ft_off, s * bx.a1, T.kmax));
        // with the Hankel amplitude b0 = +-sqrt(2 / (pi w)) folded in
        const double b0 = ti < it.count ? T.sign * sqrt(2.0 / kPi) / sqrt(w) : 0.0;
        s_fr[ti] = (cp * c0 - sp * s0) * amp * b0;
        s_fi[ti] = (sp * c0 + cp * s0) * amp * b0;
    }
    __syncthreads();
    // column chunk: all the columns up to 16 (not rounded to a power of 2:
    // rows j = (t, e2) = (j mod tb, j / tb), e2 < nsub; rows past nsub tb idle)
    const int tb = T.B < 16 ? T.B : 16;
    const int nsub = 16 / tb, G = (E + nsub - 1) / nsub;
    const int ia = lane & 15, ka = lane >> 4, jb = lane & 15, tbk = jb % tb, e2b = jb / tb;
    const int ntile = (it.count + 15) >> 4;
    // this wave's tiles u: wave + 4 u; their phi operands and window offsets
    R phiA[TPW][kIrSteps];
    int woff[TPW], nst[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int t = wave + 4 * u;
        const int b0 = 16 * t, cnt = min(16, it.count - b0);
        woff[u] = 0;
        nst[u] = 0;
#pragma unroll
        for (int st = 0; st < kIrSteps; ++st) phiA[u][st] = (R)0;
        if (t < ntile) {
            const int mw = s_m0[b0];                     // the tile's window start
            woff[u] = mw - it.mw0;
            // K steps of 4 positions that hold a kernel tap of some target
            // (the first target's taps start at mw, the last's end at its
            // m0 + kw - 1): the others are all zero and are skipped
            const int nsteps = min(kIrSteps, (s_m0[b0 + cnt - 1] + kw - mw + 3) >> 2);
            nst[u] = nsteps;
            // phi(s - m Delta) at tap k = m - m0 of the target (zero outside
            // 0 <= k < kw) from the tap's Chebyshev series in the cell offset
            // f = kw / 2 - (s - m0 Delta) / Delta
            const int mi = s_m0[b0 + ia];
            const double f = 0.5 * kw - (s_s[b0 + ia] - (double)mi * bx.delta) / bx.delta;
            R tcb[kEsN];
            cheb_basis((R)(2.0 * f - 1.0), tcb);
#pragma unroll
            for (int st = 0; st < kIrSteps; ++st) {
                const int k = mw + 4 * st + ka - mi;
                const bool on = st < nsteps && ia < cnt && k >= 0 && k < kw;
                const R v = es_tap(&s_tap[on ? k : 0][0], tcb);
                phiA[u][st] = on ? v : (R)0;
            }
        }
    }
    // the window of (column chunk cb, exponent group g) is loaded into
    // registers one group ahead (consecutive threads take consecutive
    // positions of a channel row), so that its latency overlaps the previous
    // group's products, and written to LDS between two barriers
    constexpr int PF = 16 * kIrPos / 256;
    C2 pf[PF];
    int pcp[PF];     // channel (high half) and position (low half) of this thread's r-th value
    {
        const int dc = 256 / it.npos, dp = 256 - dc * it.npos;
        int c = tid / it.npos, p = tid - c * it.npos;
#pragma unroll
        for (int r = 0; r < PF; ++r) {
            pcp[r] = (c << 16) | p;
            c += dc;
            p += dp;
            if (p >= it.npos) { p -= it.npos; ++c; }
        }
    }
    auto load_group = [&](int cb_, int g_) {
#pragma unroll
        for (int r = 0; r < PF; ++r) {
            pf[r].x = (GR)0;
            pf[r].y = (GR)0;
            if ((pcp[r] >> 16) < nsub * tb) {
                const int c = pcp[r] >> 16, p = pcp[r] & 0xffff;
                const int e = g_ * nsub + c / tb, t = cb_ + c % tb;
                if (e < E && t < T.B) {
                    int m = it.mw0 + p;
                    m = m < 0 ? m + bx.Nf : (m >= bx.Nf ? m - bx.Nf : m);
                    pf[r] = grid[bx.goff + (long long)(e * T.B + t) * bx.Nf + m];
                }
            }
        }
    };
    load_group(0, 0);
    for (int cb = 0; cb < T.B; cb += tb) {
        double Sr[TPW][4], Si[TPW][4];
#pragma unroll
        for (int u = 0; u < TPW; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) { Sr[u][v] = 0.0; Si[u][v] = 0.0; }
        for (int g = 0; g < G; ++g) {
            __syncthreads();   // the previous group's window is consumed
#pragma unroll
            for (int r = 0; r < PF; ++r)
                if ((pcp[r] >> 16) < 16) s_g[(pcp[r] >> 16) * kIrStride + (pcp[r] & 0xffff)] = pf[r];
            __syncthreads();
            if (g + 1 < G) load_group(cb, g + 1);
            else if (cb + tb < T.B) load_group(cb + tb, 0);
            const int eb = g * nsub + e2b;
            const bool eok = (cb + tbk < T.B) && e2b < nsub && eb < E;
            const bool odd = eb & 1;
            const int ec = eok ? eb : 0;
#pragma unroll
            for (int u = 0; u < TPW; ++u) {
                const int t = wave + 4 * u;
                if (t >= ntile) continue;
                const C2 *row = s_g + jb * kIrStride + woff[u] + ka;
                V dr0, di0;
#pragma unroll
                for (int v = 0; v < 4; ++v) { dr0[v] = (R)0; di0[v] = (R)0; }
                const C2 b0 = row[0], b1 = row[4], b2 = row[8];
                dr0 = Mfma<R>::mma(phiA[u][0], (R)b0.x, dr0);
                di0 = Mfma<R>::mma(phiA[u][0], (R)b0.y, di0);
                dr0 = Mfma<R>::mma(phiA[u][1], (R)b1.x, dr0);
                di0 = Mfma<R>::mma(phiA[u][1], (R)b1.y, di0);
                if (nst[u] > 2) {
                    dr0 = Mfma<R>::mma(phiA[u][2], (R)b2.x, dr0);
                    di0 = Mfma<R>::mma(phiA[u][2], (R)b2.y, di0);
                }
                if (nst[u] > 3) {
                    const C2 b3 = row[12];
                    dr0 = Mfma<R>::mma(phiA[u][3], (R)b3.x, dr0);
                    di0 = Mfma<R>::mma(phiA[u][3], (R)b3.y, di0);
                }
                if (nst[u] > 4) {
                    const C2 b4 = row[16];
                    dr0 = Mfma<R>::mma(phiA[u][4], (R)b4.x, dr0);
                    di0 = Mfma<R>::mma(phiA[u][4], (R)b4.y, di0);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 16 * t + Mfma<R>::row(lane, v);
                    const double pp = eok ? s_P[i][ec] : 0.0;
                    const double gr = (double)dr0[v];
                    const double gi = (double)di0[v];
                    // S += p i^(e mod 2) G
                    Sr[u][v] = fma(pp, odd ? -gi : gr, Sr[u][v]);
                    Si[u][v] = fma(pp, odd ? gr : gi, Si[u][v]);
                }
            }
        }
        // epilogue per tile: outputs, the exponent sub-split folded (lanes j
        // and j + tb e2 share a column), transposed through LDS so that each
        // atomic instruction adds 16 consecutive rows of 4 columns
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            const int t = wave + 4 * u;
            double out[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 16 * t + Mfma<R>::row(lane, v);
                out[v] = t < ntile ? s_fr[i] * Sr[u][v] - s_fi[i] * Si[u][v] : 0.0;
            }
            fold_exponent_split(out, tb, lane);
            // (s_out[wave] is this wave's own: wave-level ordering suffices)
            wave_lds_sync();
            if (e2b == 0) {
#pragma unroll
                for (int v = 0; v < 4; ++v) s_out[wave][Mfma<R>::row(lane, v)][jb] = out[v];
            }
            wave_lds_sync();
            if (t < ntile) {
                for (int idx = lane; idx < 16 * tb; idx += 64) {
                    const int c = idx >> 4, i = idx & 15;
                    if (16 * t + i < it.count && cb + c < T.B)
                        atomic_add_global(&T.gs[(long long)(cb + c) * T.ld_gs + bx.i0 + it.start + 16 * t + i],
                                  s_out[wave][i][c]);
                }
            }
        }
    }
}

// Interpolation, target-dense tiles, one tile per item (the form the runs
// above replace by default; knob interp_mfma = 2): 16 consecutive targets whose kernel
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
__global__ __launch_bounds__(64) void k_asy_interp_tile(const TfDesc *__restrict__ tfs,
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
    const int tb = mfma_cols(it.tb, T.B, it.col0), nsub = 16 / tb;
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
        const bool eok = colok && e2b < nsub && eb < E;
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
    fold_exponent_split(out, tb, lane);
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
            atomic_add_global(&T.gs[(long long)(cb + c) * T.ld_gs + bx.i0 + it.start + i], s_out[i][c]);
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
        atomic_add_global(&T.gs[(long long)t * T.ld_gs + bx.i0 + it.start + ti], fr * Sr - fi * Si);
}

}  // namespace nufht_dev
