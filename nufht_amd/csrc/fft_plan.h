// Launch geometry of the two-pass FFT kernels (fft.h), shared with the host
// code that plans the FFT size classes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace nufht_dev {

constexpr int kFftThreads = 512;
// M * N per block: 64 KiB of data, so that two blocks share a CU and one
// block's loads overlap the other's LDS passes (the kernels are bounded to
// 128 VGPRs, 4 waves per SIMD, for the same reason)
constexpr int kFftTileBytes = 65536;
template <typename R> constexpr int fft_max_elems() { return kFftTileBytes / (2 * (int)sizeof(R)); }
constexpr int kFftMaxStages = 8;

struct FftPass {
    int N, M;                 // transform length; sequences per block (a power of 2)
    int lgM, lgN;             // log2 M; log2 N when N is a power of 2
    int N1, N2;               // the class's factorisation, Nf = N1 * N2
    int tiles;                // blocks per channel (columns or rows / M)
    int nstage;
    int radix[kFftMaxStages]; // powers of 2 first, then at most one 3 or 9
    int w_off;                // W_N[m] = e^{2 pi i m / N}, m < N
    int hi_off, lo_off;       // pass 1: w^a = hi[a >> 8] lo[a & 255], a < Nf
    int lds_stride, lds_shift; // LDS layout: fft_lds_idx
    int chinfo_off;           // per channel {L, mlo, span}: pass 1 reads only grid positions
                              // l in [-L, L] (the rest are zeros not stored), pass 2
                              // writes only frequencies mlo .. mlo + span (mod Nf)
    long long nf;
    int nch;                  // single-pass classes: channels of the class
    int pad_;
};

__host__ __device__ inline size_t fft_lds_elems(int N, int stride, int shift, int M)
{
    return (size_t)N * stride + (N >> shift) + M;
}

}  // namespace nufht_dev
