// Spread and interpolation launchers of the asymptotic regime for one grid
// precision and exponent bound, selected at build time:
//   -DNUFHT_ASY_R=double|float -DNUFHT_ASY_EM=6|10|14|22
// (each pair is its own object, so that the kernels' instantiations compile
// in parallel).
#include "kernels_asy.h"
#include "kernels_fused.h"
#include "program.h"

#include <mutex>

#ifndef NUFHT_ASY_R
#error "NUFHT_ASY_R must name the grid precision (double or float)"
#endif
#ifndef NUFHT_ASY_EM
#error "NUFHT_ASY_EM must give the exponent bound (6, 10, 14 or 22)"
#endif

namespace nufht_dev {

// The matrix-core spread's launches, one per products-per-wave count KW
// (items [adense_off[KW], adense_off[KW + 1]) of the program's list).
template <typename R, int EM, bool SP, int KW>
static void spread_mfma_launch(Program *P, hipStream_t st)
{
    typedef typename Cplx<R>::T G;
    if constexpr (KW >= 1) {
        const int a = P->adense_off[KW], b = P->adense_off[KW + 1];
        if (b > a) {
            const size_t lds = SpreadSched<SP>::template lds<R, EM>();
            static std::once_flag attr_once;
            std::call_once(attr_once, [lds]() {
                hipFuncSetAttribute((const void *)k_asy_spread_mfma<R, EM, SP, KW>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            });
            hipLaunchKernelGGL((k_asy_spread_mfma<R, EM, SP, KW>), dim3((unsigned)(b - a)),
                               dim3(SpreadSched<SP>::threads), lds, st, P->d_tfs, P->d_abox,
                               P->d_adense + a, P->d_adrange + a, (G *)P->d_grid, P->d_ft);
        }
        spread_mfma_launch<R, EM, SP, KW - 1>(P, st);
    }
}

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
        if (P->spread_split) spread_mfma_launch<R, EM, true, (2 * EM + 3) / 4>(P, st);
        else spread_mfma_launch<R, EM, false, (2 * EM + 3) / 4>(P, st);
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

// (internal linkage: every exponent-bound object defines its own copy)
static __global__ void k_queue_reset(int *q) { *q = 0; }

// The fused small boxes: one persistent launch over the unit queue (its
// counter cleared by a kernel first, so that the pair can be captured in a
// graph and replayed).
template <typename R, int EM>
void asy_fused(Program *P, hipStream_t st)
{
    typedef typename Cplx<R>::T G;
    if (P->funits.empty()) return;
    StageTimer tm(ST_FUSED, st);
    const size_t lds = (size_t)P->fused_grid_elems * sizeof(G) + fused_stage_bytes<R, EM>();
    static std::once_flag attr_once;
    std::call_once(attr_once, []() {
        hipFuncSetAttribute((const void *)k_asy_fused<R, EM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kFusedGridBytes + fused_stage_bytes<R, EM>());
    });
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, st, P->d_fqueue);
    hipLaunchKernelGGL((k_asy_fused<R, EM>), dim3((unsigned)P->fused_blocks), dim3(kFusedThreads), lds,
                       st, P->d_tfs, P->d_abox, P->d_funits, P->d_fpass, (int)P->funits.size(),
                       P->d_fqueue, (const G *)P->d_fftw, P->d_ft, P->fused_grid_elems);
}

template void asy_fused<NUFHT_ASY_R, NUFHT_ASY_EM>(Program *, hipStream_t);
template void asy_spread<NUFHT_ASY_R, NUFHT_ASY_EM>(Program *, hipStream_t);
template void asy_interp<NUFHT_ASY_R, NUFHT_ASY_EM>(Program *, hipStream_t);

}  // namespace nufht_dev
