// Grid zeroing and uniform-FFT launchers of the asymptotic regime for one
// grid precision, selected at build time: -DNUFHT_ASY_R=double|float.
// Classes the two-pass FFT kernels cover run them (fft.h); the rest run
// rocFFT out of place.
#include "fft.h"
#include "kernels_asy.h"
#include "program.h"

#include <mutex>

#ifndef NUFHT_ASY_R
#error "NUFHT_ASY_R must name the grid precision (double or float)"
#endif

namespace nufht_dev {

template <typename R>
void asy_grid_zero(Program *P, hipStream_t st)
{
    typedef typename Cplx<R>::T G;
    if (P->azero.empty()) return;
    StageTimer tm(ST_GZERO, st);
    hipLaunchKernelGGL((k_grid_zero<R>), dim3((unsigned)P->azero.size()), dim3(256), 0, st,
                       P->d_tfs, P->d_abox, P->d_azero, (G *)P->d_grid);
}

template <typename R>
void asy_fft(Program *P, hipStream_t st)
{
    typedef typename Cplx<R>::T G;
    static std::once_flag attr_once;
    std::call_once(attr_once, []() {
        hipFuncSetAttribute((const void *)k_fft_cols<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
        hipFuncSetAttribute((const void *)k_fft_rows<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
        hipFuncSetAttribute((const void *)k_fft_small<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
    });
    StageTimer tm(ST_FFT, st);
    for (auto &c : P->fft) {
        G *in = (G *)P->d_grid + c.off, *out = (G *)P->d_grid2 + c.off;
        if (c.own && c.small) {
            hipLaunchKernelGGL((k_fft_small<R>), dim3((unsigned)((c.batch + c.p1.M - 1) / c.p1.M)),
                               dim3(kFftThreads), c.lds1, st, (const G *)in, out, c.p1,
                               (const G *)P->d_fftw, (const int4 *)P->d_fft_chinfo);
            continue;
        }
        if (c.own) {
            const G *tw = (const G *)P->d_fftw;
            hipLaunchKernelGGL((k_fft_cols<R>), dim3((unsigned)(c.batch * c.p1.tiles)),
                               dim3(kFftThreads), c.lds1, st, in, c.p1, tw,
                               (const int4 *)P->d_fft_chinfo);
            hipLaunchKernelGGL((k_fft_rows<R>), dim3((unsigned)(c.batch * c.p2.tiles)),
                               dim3(kFftThreads), c.lds2, st, (const G *)in, out, c.p2, tw,
                               (const int4 *)P->d_fft_chinfo);
            continue;
        }
        rocfft_execution_info_set_stream(c.info, st);
        void *ib = (void *)in, *ob = (void *)out;
        rocfft_execute(c.plan, &ib, &ob, c.info);
    }
}

template void asy_grid_zero<NUFHT_ASY_R>(Program *, hipStream_t);
template void asy_fft<NUFHT_ASY_R>(Program *, hipStream_t);

}  // namespace nufht_dev
