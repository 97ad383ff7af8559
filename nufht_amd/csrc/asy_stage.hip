// Spread and interpolation launchers of the asymptotic regime for one grid
// precision and exponent bound, selected at build time:
//   -DNUFHT_ASY_R=double|float -DNUFHT_ASY_EM=6|10|14|22
// (each pair is its own object, so that the kernels' instantiations compile
// in parallel).
#include "kernels_asy.h"
#include "program.h"

#include <mutex>

#ifndef NUFHT_ASY_R
#error "NUFHT_ASY_R must name the grid precision (double or float)"
#endif
#ifndef NUFHT_ASY_EM
#error "NUFHT_ASY_EM must give the exponent bound (6, 10, 14 or 22)"
#endif

namespace nufht_dev {

template <typename R, int EM>
void asy_spread(Program *P, hipStream_t st)
{
    typedef typename Cplx<R>::T G;
    if (!P->aspread.empty()) {
        StageTimer tm(ST_SPREAD, st);
        hipLaunchKernelGGL((k_asy_spread<R, EM>), dim3((unsigned)P->aspread.size()), dim3(256),
                           0, st, P->d_tfs, P->d_abox, P->d_aspread, P->d_arange,
                           (G *)P->d_grid, P->d_ft);
    }
    if (!P->adense.empty()) {
        StageTimer tm(ST_SPREAD_MF, st);
        static std::once_flag attr_once;
        std::call_once(attr_once, []() {
            hipFuncSetAttribute((const void *)k_asy_spread_mfma<R, EM, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)SpreadSched<true>::template lds<R, EM>());
            hipFuncSetAttribute((const void *)k_asy_spread_mfma<R, EM, false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)SpreadSched<false>::template lds<R, EM>());
        });
        const size_t lds_split = SpreadSched<true>::template lds<R, EM>();
        const size_t lds_unified = SpreadSched<false>::template lds<R, EM>();
        if (P->spread_split)
            hipLaunchKernelGGL((k_asy_spread_mfma<R, EM, true>), dim3((unsigned)P->adense.size()),
                               dim3(SpreadSched<true>::threads), lds_split, st, P->d_tfs,
                               P->d_abox, P->d_adense, P->d_adrange, (G *)P->d_grid, P->d_ft);
        else
            hipLaunchKernelGGL((k_asy_spread_mfma<R, EM, false>), dim3((unsigned)P->adense.size()),
                               dim3(SpreadSched<false>::threads), lds_unified, st, P->d_tfs,
                               P->d_abox, P->d_adense, P->d_adrange, (G *)P->d_grid, P->d_ft);
    }
}

template <typename R, int EM>
void asy_interp(Program *P, hipStream_t st)
{
    typedef typename Cplx<R>::T G;
    if (!P->ainterp.empty()) {
        StageTimer tm(ST_INTERP, st);
        int tmx = P->interp_tmax;
        size_t lds = interp_lds_off(tmx, 2, sizeof(R)) * sizeof(double);
        hipLaunchKernelGGL((k_asy_interp<R, EM>), dim3((unsigned)P->ainterp.size()), dim3(256),
                           lds, st, P->d_tfs, P->d_abox, P->d_ainterp, (const G *)P->d_grid2,
                           P->d_ft, tmx);
    }
    if (!P->aimf.empty()) {
        StageTimer tm(ST_INTERP_MF, st);
        // the products in the grid's precision: FP32 matrix cores (twice the
        // FP64 rate) for the FP32 grids of mixed precision
        hipLaunchKernelGGL((k_asy_interp_mfma<R, R, EM>), dim3((unsigned)P->aimf.size()), dim3(256),
                           0, st, P->d_tfs, P->d_abox, P->d_aimf, (const G *)P->d_grid2, P->d_ft);
    }
    if (!P->aimt.empty()) {
        StageTimer tm(ST_INTERP_MF, st);
        hipLaunchKernelGGL((k_asy_interp_tile<R, double, EM>), dim3((unsigned)P->aimt.size()),
                           dim3(64), 0, st, P->d_tfs, P->d_abox, P->d_aimt, (const G *)P->d_grid2,
                           P->d_ft);
    }
}

template void asy_spread<NUFHT_ASY_R, NUFHT_ASY_EM>(Program *, hipStream_t);
template void asy_interp<NUFHT_ASY_R, NUFHT_ASY_EM>(Program *, hipStream_t);

}  // namespace nufht_dev
