// Definitions shared by every device translation unit of the MI355X NUFHT
// engine: work-item and transform descriptors, the kernel-transform and
// tap tables' evaluators, fast phases, complex and matrix-core helpers.
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
// targets [start, start + count) of an asymptotic box, interpolated from the
// grid positions [mw0, mw0 + npos)
struct InterpRun { int box, start, count, mw0, npos, pad; };
constexpr int kIrTiles = 8;    // 16-target tiles per matrix-core interpolation run
constexpr int kIrPos = 64;     // grid positions a run stages, at most
constexpr int kIrSteps = 5;    // K steps of 4 positions per tile: windows of up to 20
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
// quadrant of n.  Arguments beyond 2^21 go to the library routine: inlined
// (SC_INLINE), out of line (SC_CALL: its full-range reduction holds about 40
// registers wherever it is inlined, although only |v| >= 2^21 reaches it), or
// not at all (SC_REDUCED: the caller guarantees |v| < 2^21).
enum { SC_INLINE, SC_CALL, SC_REDUCED };
__device__ __attribute__((noinline)) inline void sincos_full(double v, double *s, double *c)
{
    sincos(v, s, c);
}

template <int MODE = SC_INLINE>
__device__ __forceinline__ void sincos_cw(double v, double *s, double *c)
{
    if (MODE != SC_REDUCED && !(fabs(v) < 2097152.0)) {
        if (MODE == SC_CALL) sincos_full(v, s, c);
        else sincos(v, s, c);
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

// sph_jn (bessel_dev.h) with the Cody-Waite reduction above (its full-range
// fallback out of line) instead of the library's sincos: the local
// expansions' arguments stay below the split point z_split, far inside the
// range where the reduction is exact to an ulp.
__device__ inline double sph_jn_cw(int n, double hi, double lo)
{
    const double x = hi;
    if (x == 0.0) return n == 0 ? 1.0 : 0.0;
    if (x <= fmax(2.0, (double)n)) return series_j(n, 1, x) * sqrt(kPi / (2.0 * x));
    double sh, ch;
    sincos_cw<SC_CALL>(hi, &sh, &ch);
    const double s = fma(lo, ch, sh), c = fma(-lo, sh, ch);
    const double xi = 1.0 / x;
    double j0 = s * xi;
    if (n == 0) return j0;
    double j1 = (s * xi - c) * xi;
    for (int k = 1; k < n; ++k) {
        const double j2 = (double)(2 * k + 1) * xi * j1 - j0;
        j0 = j1;
        j1 = j2;
    }
    return j1;
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

// A load through a global-address-space pointer.  Pointers read from a
// descriptor in memory are generic, and generic loads (flat) count against
// the LDS / scalar wait counter as well, so that a later wait for an LDS
// read or a scalar load also waits for them; global loads count only
// against the vector-memory counter, and loads issued ahead stay in flight.
template <typename T>
__device__ __forceinline__ T ld_global(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}

// An FP64 atomic add into global memory through a global-address-space
// pointer (a global atomic, not a flat one: see ld_global).
__device__ __forceinline__ void atomic_add_global(double *p, double v)
{
    __hip_atomic_fetch_add((__attribute__((address_space(1))) double *)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Orders one wave's LDS writes before its later LDS reads (and reads before
// later writes) where only that wave touches the data: no block barrier.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Matrix-core interpolation epilogue: the 16 lanes jb = lane mod 16 of each
// quarter wave hold rows j = (t, e2) = (jb mod tb, jb / tb), e2 < nsub = 16
// / tb; sums the nsub parts of each column t into lane jb = t (the other
// lanes' results are not used).  tb need not be a power of 2.
__device__ __forceinline__ void fold_exponent_split(double (&out)[4], int tb, int lane)
{
    if ((tb & (tb - 1)) == 0) {
        for (int off = tb; off < 16; off <<= 1)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[v] += __shfl_xor(out[v], off, 64);
        return;
    }
    const int nsub = 16 / tb, jb = lane & 15;
    double o0[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) o0[v] = out[v];
    for (int q = 1; q < nsub; ++q) {
        const int src = (lane & ~15) | min(jb + q * tb, 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) out[v] += __shfl(o0[v], src, 64);
    }
}

constexpr int kVQ = 64;            // sources staged per LDS pass of the VALU spread
// Exponents per lane of the VALU spread (4 EG accumulators: two columns,
// complex): channels are split over lanes, so a lane's registers do not grow
// with the box's E.
template <typename R> struct SpreadEG { static constexpr int v = 4; };
template <> struct SpreadEG<float> { static constexpr int v = 8; };
// lanes per grid position of a VALU spread item: column pairs x exponent groups
__host__ __device__ inline int spread_valu_lanes(int E, int tb, bool f32)
{
    const int eg = f32 ? 8 : 4;
    return (tb > 1 ? tb >> 1 : 1) * ((E + eg - 1) / eg);
}
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

constexpr int kSQ = 64;           // sources staged per LDS pass
constexpr int kSP = 16;           // accumulator columns: a window of grid positions

// Sources [piece.x, piece.y) of a box; the item writes grid positions l in
// [wr.x, wr.y], with plain stores in [ex.x, ex.y] (no other item reaches
// them) and atomic adds elsewhere (zeroed beforehand).
struct SpreadRange { int2 piece, wr, ex, pad; };

// Columns an interpolation tile's 16 rows hold: its column chunk (tb, a
// power of 2) cut to the columns present, so that rows j = (t, e2) = (j mod
// tbm, j / tbm) carry nsub = 16 / tbm exponents (5 columns: 3 exponents per
// group rather than 2); rows j >= nsub tbm are idle.
__host__ __device__ inline int mfma_cols(int tb, int B, int col0)
{
    return tb < B - col0 ? tb : B - col0;
}
// Products per consumer wave of an item: 2 per exponent group (real and
// imaginary part), nsub = 16 / tb exponents per group, split over 4 waves.
__host__ __device__ inline int spread_mfma_kw(int E, int tb)
{
    const int nsub = 16 / tb;
    return (2 * ((E + nsub - 1) / nsub) + 3) >> 2;
}

}  // namespace nufht_dev
