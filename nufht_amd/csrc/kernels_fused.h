// Small asymptotic boxes in one persistent launch: spread, uniform FFT and
// interpolation of a box's channels with the uniform grid held in LDS.
//
// Most asymptotic boxes have small uniform grids (Nf <= 256 points, 2L + 1
// spread positions) but many channels (C = E B): through the staged pipeline
// each such box writes its grid to HBM, waits for an FFT launch of its size
// class and is read back by the interpolation launch, and its per-source and
// per-target set-up is spread over three kernels' items.  Here one block
// takes a unit (box, column chunk of tb columns, exponent range [e0, e0 +
// ne)), whose ne tb channels of Nf points fit in LDS, and runs
//
//   spread:  H_l[c] = dec(l) sum_j v_j[c] psi(l h - x_j)     (VALU, into LDS)
//   FFT:     T_m[c] = sum_l H_l[c] e^{2 pi i l m / Nf}         (Stockham, in LDS)
//   interp:  g[i, t] += Re( f_i sum_e P_e(w_i) i^(e mod 2)
//                          sum_k phi_k(s_i) T_{m_i + k}[(e, t)] )   (matrix cores)
//
// with the quantities of the staged kernels (kernels_asy.h): v_j[(e, t)] =
// c_j,t (rho / r_j)^e r_j^{-1/2} e^{i D x_j}, dec the kernel transform's
// reciprocal, f_i the phase, deconvolution and weight of target i.  Blocks
// are persistent: each takes units from a work queue (an atomic counter over
// the unit list, which the host orders by decreasing cost) until it is
// drained, so that every box of every transform in a program runs in this
// one launch, large units first.
//
// Measured against the staged kernels on the benchmark workloads (uneven
// 256-transform batch, headline, clustered): the staged kernels are faster
// for every box size tried except the smallest (at most 128 sources and
// targets), because a unit holds at most 54 KiB of grid, so boxes with many
// channels are cut into several units that each repeat the per-source and
// per-target set-up.  The path is therefore off by default (knob fused_max)
// and the autotuner tries it.
//
// The three phases are separate (not inlined) functions so that each gets
// its own register allocation; at 256 VGPRs (one 512-thread block per CU)
// none of them spills.
#pragma once
#include "device_common.h"
#include "fft.h"

namespace nufht_dev {

constexpr int kFusedThreads = kFftThreads;   // 512: the in-LDS FFT's block
constexpr int kFQ = 64;                       // sources staged per pass of the spread
constexpr int kFusedNfMax = 256;              // uniform grid points of a fused box, at most
constexpr int kFusedGridBytes = 54 * 1024;    // LDS of a unit's grid, at most

// Complex channels a spread lane accumulates (CB columns x EB exponents).
__host__ __device__ constexpr int fused_lane_ch(bool f32) { return f32 ? 16 : 8; }

// Unit: columns [col0, col0 + tb), exponents [e0, e0 + ne) of asymptotic box
// `box`; pass `fp` describes its grid (Nf points, M >= ne tb channels, LDS layout
// and twiddles).
struct FusedUnit { int box, col0, e0, ne, tb, fp; };

// Staging area of the spread (after the grid in LDS), in bytes.
template <typename R, int EMAX> __host__ __device__ constexpr int fused_stage_bytes()
{
    constexpr int NG = (EMAX + fused_lane_ch(sizeof(R) == 4) / 2 - 1) /
                       (fused_lane_ch(sizeof(R) == 4) / 2);
    return kFQ * (8 + 16) + (kFusedNfMax / 2) * 8 + (int)sizeof(R) * kFQ * (kTapMax + 16 + NG + 1) +
           4 * kFQ;
}

// x^n for n >= 0 by repeated squaring
__device__ __forceinline__ double pow_int(double x, int n)
{
    double r = 1.0;
    while (n) {
        if (n & 1) r *= x;
        x *= x;
        n >>= 1;
    }
    return r;
}

// The spread of one unit, added into the zeroed LDS grid s (layout
// fft_lds_idx).  Lane: a position l (positions fastest), CB columns, EB
// exponents and one of K interleaved subsets of the sources (K: as many as
// the lanes left over by positions x channel groups allow), so that small
// grids keep the block's lanes busy; each lane adds its partial sums into the
// grid at the end.
template <typename R, int EMAX, int CB, int EB>
__device__ __attribute__((noinline)) void fused_spread(const TfDesc &T, const AsyBoxD &bx,
                                                       const FusedUnit &u, const FftPass &fp,
                                                       typename Cplx<R>::T *s, unsigned char *stg,
                                                       const double *__restrict__ ft)
{
    constexpr int NG = (EMAX + fused_lane_ch(sizeof(R) == 4) / 2 - 1) /
                       (fused_lane_ch(sizeof(R) == 4) / 2);   // groups at EB = lane_ch / 2
    double *s_x = (double *)stg;
    double2 *s_ph = (double2 *)(s_x + kFQ);
    double *s_dec = (double *)(s_ph + kFQ);
    R(*s_psi)[kTapMax] = (R(*)[kTapMax])(s_dec + kFusedNfMax / 2);
    R(*s_c)[kFQ] = (R(*)[kFQ])(&s_psi[kFQ][0]);                 // [column][source]
    R(*s_pw)[NG + 1] = (R(*)[NG + 1])(&s_c[16][0]);            // [source][group]; NG: rho / r
    int *s_pf = (int *)(&s_pw[kFQ][0]);
    const int tid = threadIdx.x;
    const int P = 2 * bx.L + 1;
    const int ntl = u.tb / CB, ng = (u.ne + EB - 1) / EB, lpp = ntl * ng;
    const int K = max(1, kFusedThreads / (P * lpp));           // source subsets
    const int pl = tid % P, rest = tid / P;
    const int combo = rest % lpp, ks = rest / lpp;
    const int tl = combo % ntl, g = combo / ntl;
    const bool on = ks < K;                 // the host keeps P lpp <= blockDim
    const int l = pl - bx.L;
    const int kw = T.kw;
    const double inv_a1 = 1.0 / bx.a1;
    R ar[CB][EB], ai[CB][EB];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int e = 0; e < EB; ++e) { ar[c][e] = (R)0; ai[c][e] = (R)0; }
    for (int c0 = bx.j0; c0 < bx.j0 + bx.nj; c0 += kFQ) {
        const int cn = min(kFQ, bx.j0 + bx.nj - c0);
        if (tid < cn) {
            const double rj = T.r[c0 + tid];
            const double x = rj - bx.C;
            s_x[tid] = x;
            s_pf[tid] = (int)ceil((x - bx.a1) / bx.h);
            double sn, cs;
            sincos_cw<SC_REDUCED>(bx.D * x, &sn, &cs);
            const double rs = 1.0 / sqrt(rj);
            s_ph[tid] = make_double2(cs * rs, sn * rs);
            const double ri = bx.rho / rj;
            const double rg = pow_int(ri, EB);
            double pw = pow_int(ri, u.e0);
#pragma unroll
            for (int q = 0; q < NG; ++q) { s_pw[tid][q] = (R)pw; pw *= rg; }
            s_pw[tid][NG] = (R)ri;
        }
        // coefficients: consecutive lanes on consecutive sources of a column
        for (int q = tid; q < u.tb * cn; q += kFusedThreads) {
            const int t = q / cn, jq = q - t * cn;
            const int tc = u.col0 + t;
            s_c[t][jq] = tc < T.B ? (R)T.cs[(long long)tc * T.ld_cs + c0 + jq] : (R)0;
        }
        __syncthreads();
        for (int q = tid; q < cn * (kw + 1); q += kFusedThreads) {
            const int jq = q / (kw + 1), k = q - jq * (kw + 1);
            const double uu = ((double)(s_pf[jq] + k) * bx.h - s_x[jq]) * inv_a1;
            s_psi[jq][k] = (R)es_eval((R)uu, (R)T.beta);
        }
        __syncthreads();
        if (on) {
            // sources of the chunk whose taps reach l: lf in [l - kw, l] (lf ascending)
            int a = 0, b = cn;
            while (a < b) { int mid = (a + b) >> 1; if (s_pf[mid] >= l - kw) b = mid; else a = mid + 1; }
            const int qa = a;
            b = cn;
            while (a < b) { int mid = (a + b) >> 1; if (s_pf[mid] > l) b = mid; else a = mid + 1; }
            const int qb = a;
            for (int q = qa + ks; q < qb; q += K) {
                const R w = s_psi[q][l - s_pf[q]] * s_pw[q][g];
                const double2 ph = s_ph[q];
                const R zr = (R)ph.x * w, zi = (R)ph.y * w, ri = s_pw[q][NG];
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    const R cv = s_c[tl * CB + c][q];
                    const R vr = zr * cv, vi = zi * cv;
                    R rp = (R)1;
#pragma unroll
                    for (int e = 0; e < EB; ++e) {
                        ar[c][e] = fma(vr, rp, ar[c][e]);
                        ai[c][e] = fma(vi, rp, ai[c][e]);
                        rp *= ri;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid < P)
        s_dec[tid] = 1.0 / (bx.a2 * ft_eval(ft + T.ft_off, (double)(tid - bx.L) * bx.h * bx.a2, T.kmax));
    __syncthreads();
    if (on) {
        const double dec = s_dec[pl];
        const int n = l < 0 ? l + bx.Nf : l;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const int t = tl * CB + c;
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                const int el = g * EB + e;
                if (el < u.ne && t < u.tb) {
                    R *d = (R *)&s[fft_lds_idx(n, el * u.tb + t, fp.lds_stride, fp.lds_shift)];
                    if (K == 1) {
                        d[0] = (R)((double)ar[c][e] * dec);
                        d[1] = (R)((double)ai[c][e] * dec);
                    } else {
                        atomicAdd(d, (R)((double)ar[c][e] * dec));
                        atomicAdd(d + 1, (R)((double)ai[c][e] * dec));
                    }
                }
            }
        }
    }
    __syncthreads();
}

// Interpolation of every target of the unit's box from the transformed LDS
// grid: 16-target tiles, one per wave at a time, on the matrix cores (rows
// the tile's targets, K the window's positions, columns (exponent sub-split
// e2, column t) as in k_asy_interp_mfma); the unit's share of each output is
// added to g.
template <typename R>
__device__ __attribute__((noinline)) void fused_interp(const TfDesc &T, const AsyBoxD &bx, const FusedUnit &u,
                                             const FftPass &fp, const typename Cplx<R>::T *s,
                                             unsigned char *stg, const double *__restrict__ ft)
{
    typedef typename Cplx<R>::T C2;
    typedef typename Mfma<R>::V V;
    constexpr int KS = 4;                     // K steps (4 positions each) per pass over the groups
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nwave = kFusedThreads / 64;
    // per wave: the tile's targets (s, fr, fi, bw, rw; m0)
    double *w_s = (double *)stg + wave * 96;
    double *w_fr = w_s + 16, *w_fi = w_s + 32, *w_bw = w_s + 48, *w_rw = w_s + 64;
    int *w_m0 = (int *)(w_s + 80);
    const int tb = u.tb, nsub = 16 / tb, G = (u.ne + nsub - 1) / nsub;
    const int ia = lane & 15, ka = lane >> 4, jb = lane & 15, tbk = jb % tb, e2b = jb / tb;
    const int kw = T.kw, Nf = bx.Nf;
    const bool tok = u.col0 + tbk < T.B;
    const int ntile = (bx.ni + 15) >> 4;
    for (int tile = wave; tile < ntile; tile += nwave) {
        const int b0 = 16 * tile, cnt = min(16, bx.ni - b0);
        if (lane < 16) {
            // target lane of the tile (rows past the end repeat the last with zero weight)
            const double w = T.w[bx.i0 + b0 + min(lane, cnt - 1)];
            const double sv = w - bx.D;
            const double nu = (double)T.order + 0.5 * (double)T.half;
            const int qph = (int)lround(2.0 * nu + 1.0) & 7;
            const double c0 = cos_qpi4(qph), s0 = -cos_qpi4((qph + 6) & 7);
            double sp, cp;
            sincos_cw<SC_REDUCED>(w * bx.C, &sp, &cp);
            const double amp = bx.h * bx.delta / (bx.a1 * ft_eval(ft + T.ft_off, sv * bx.a1, T.kmax));
            w_s[lane] = sv;
            w_m0[lane] = (int)ceil((sv - bx.a2) / bx.delta);
            w_fr[lane] = (cp * c0 - sp * s0) * amp;
            w_fi[lane] = (sp * c0 + cp * s0) * amp;
            w_bw[lane] = lane < cnt ? T.sign * sqrt(2.0 / kPi) / sqrt(w) : 0.0;
            w_rw[lane] = 1.0 / (bx.rho * w);
        }
        // the wave's own LDS writes are visible to its later reads (in-order
        // LDS per wave); keep the compiler from moving them across
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const double sv = w_s[ia];
        const int mw = w_m0[0] - 1;                            // the tile's window start
        const int nst = (w_m0[cnt - 1] + kw - mw + 3) >> 2;
        double pw0[4], rwn[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double rv = w_rw[Mfma<R>::row(lane, v)];
            double p = 1.0, pn = 1.0;
            for (int e = 0; e < u.e0 + e2b; ++e) p *= rv;
            for (int e = 0; e < nsub; ++e) pn *= rv;
            pw0[v] = p;          // rw^e for this lane's exponent of group 0
            rwn[v] = pn;         // rw^nsub: one group further
        }
        double Sr[4] = {0.0, 0.0, 0.0, 0.0}, Si[4] = {0.0, 0.0, 0.0, 0.0};
        // S is linear in the window's positions: passes of KS K steps
        for (int st0 = 0; st0 < nst; st0 += KS) {
            R phiA[KS];
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                const double x = (sv - (double)(mw + 4 * (st0 + k) + ka) * bx.delta) / bx.a2;
                phiA[k] = (st0 + k < nst && ia < cnt) ? (R)es_kernel(x, T.beta) : (R)0;
            }
            double pwv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) pwv[v] = pw0[v];
            for (int gq = 0; gq < G; ++gq) {
                const int el = gq * nsub + e2b, e = u.e0 + el;
                const bool eok = tok && el < u.ne;
                const int ch = (eok ? el : 0) * tb + tbk;
                V dr, di;
#pragma unroll
                for (int v = 0; v < 4; ++v) { dr[v] = (R)0; di[v] = (R)0; }
#pragma unroll
                for (int k = 0; k < KS; ++k) {
                    if (st0 + k < nst) {
                        int m = mw + 4 * (st0 + k) + ka;
                        m = ((m % Nf) + Nf) % Nf;
                        const C2 bv = s[fft_lds_idx(m, ch, fp.lds_stride, fp.lds_shift)];
                        dr = Mfma<R>::mma(phiA[k], eok ? bv.x : (R)0, dr);
                        di = Mfma<R>::mma(phiA[k], eok ? bv.y : (R)0, di);
                    }
                }
                const double sg = eok ? ((((e >> 1) & 1) ? -T.a[e] : T.a[e])) : 0.0;
                const bool odd = e & 1;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double pp = sg * pwv[v];
                    const double gr = (double)dr[v], gi = (double)di[v];
                    // S += P_e i^(e mod 2) G
                    Sr[v] = fma(pp, odd ? -gi : gr, Sr[v]);
                    Si[v] = fma(pp, odd ? gr : gi, Si[v]);
                    pwv[v] *= rwn[v];
                }
            }
        }
        double out[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = Mfma<R>::row(lane, v);
            out[v] = w_bw[i] * (w_fr[i] * Sr[v] - w_fi[i] * Si[v]);
        }
        // lanes j and j + tb e2 hold the same column: fold the sub-split
        for (int off = tb; off < 16; off <<= 1)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[v] += __shfl_xor(out[v], off, 64);
        if (e2b == 0 && tok) {
            double *gcol = T.gs + (long long)(u.col0 + tbk) * T.ld_gs + bx.i0 + b0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = Mfma<R>::row(lane, v);
                if (i < cnt) atomicAdd(gcol + i, out[v]);
            }
        }
        // the next tile's writes must not overtake this tile's reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <typename R>
__device__ __attribute__((noinline)) void fused_fft(typename Cplx<R>::T *s, const FftPass &fp, const typename Cplx<R>::T *tw)
{
    fft_lds<R>(s, fp, tw + fp.w_off);
}

template <typename R, int EMAX>
__global__ __launch_bounds__(kFusedThreads, 2) void k_asy_fused(const TfDesc *__restrict__ tfs,
                                                                 const AsyBoxD *__restrict__ boxes,
                                                                 const FusedUnit *__restrict__ units,
                                                                 const FftPass *__restrict__ passes,
                                                                 int nunits, int *__restrict__ queue,
                                                                 const typename Cplx<R>::T *__restrict__ tw,
                                                                 const double *__restrict__ ft,
                                                                 int grid_elems)
{
    typedef typename Cplx<R>::T C2;
    extern __shared__ double s_fused[];
    C2 *s = (C2 *)s_fused;
    unsigned char *stg = (unsigned char *)(s + grid_elems);
    __shared__ int s_unit;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(queue, 1);
        __syncthreads();
        const int ui = s_unit;
        __syncthreads();
        if (ui >= nunits) break;
        const FusedUnit u = units[ui];
        const AsyBoxD bx = boxes[u.box];
        const TfDesc &T = tfs[bx.tf];
        const FftPass fp = passes[u.fp];
        const int ne = (int)fft_lds_elems(fp.N, fp.lds_stride, fp.lds_shift, fp.M);
        for (int q = threadIdx.x; q < ne; q += kFusedThreads) {
            C2 z;
            z.x = (R)0;
            z.y = (R)0;
            s[q] = z;
        }
        __syncthreads();
        if (u.tb == 1)
            fused_spread<R, EMAX, 1, fused_lane_ch(sizeof(R) == 4)>(T, bx, u, fp, s, stg, ft);
        else
            fused_spread<R, EMAX, 2, fused_lane_ch(sizeof(R) == 4) / 2>(T, bx, u, fp, s, stg, ft);
        fused_fft<R>(s, fp, tw);
        fused_interp<R>(T, bx, u, fp, s, stg, ft);
        __syncthreads();
    }
}

}  // namespace nufht_dev
