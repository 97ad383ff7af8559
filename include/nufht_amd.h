/**
 * \file nufht_amd.h
 * \brief Extensions of the NUFHT C ABI for device-resident data, adjoints,
 *        node gradients, uneven multi-transform batches, precision policy,
 *        per-stage timing and autotuning.  None of these has a counterpart
 *        in the reference; each says which reference behaviour it builds on.
 *
 * Streams are passed as `void *` holding a hipStream_t (NULL = the null
 * stream).  Device pointers are HIP device memory (e.g. torch CUDA tensors
 * on ROCm).  Column-major layouts with leading dimensions, as in
 * fast_hankel_nufht.h.
 */
#ifndef NUFHT_AMD_H
#define NUFHT_AMD_H

#include "fast_hankel_nufht.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- precision policy ---------------------------------------------------
 * FP64: every stage in double precision (the reference's arithmetic).
 * AUTO (default): when tol >= 1e-6, spreading, interpolation, the uniform
 *   FFT and the local-expansion contraction run in FP32 with FP32
 *   accumulation; phases, deconvolution, Bessel evaluation and expansion
 *   coefficients stay FP64.  tol < 1e-6 runs FP64.
 * MIXED: the FP32 stages regardless of tol (for experiments). */
enum { NUFHT_PREC_FP64 = 0, NUFHT_PREC_AUTO = 1, NUFHT_PREC_MIXED = 2 };
int nufht_amd_plan_set_precision(NufhtPlan *plan, int mode);
/** Effective precision of the plan's execution: 0 = FP64, 2 = mixed. */
int nufht_amd_plan_precision(const NufhtPlan *plan);

/* ---- device-resident execution ------------------------------------------
 * Same transform as nufht_batch (reference src/fast_hankel_nufht.c:989),
 * with cs/gs in device memory, enqueued on `stream` (asynchronous). */
int nufht_batch_device(const NufhtPlan *plan,
                       const double *d_cs, int ld_cs,
                       double *d_gs, int ld_gs,
                       int batch, void *stream);

/* ---- adjoint ---------------------------------------------------------------
 * x[j,t] = sum_i y[i,t] J_nu(ws_i rs_j): the transposed transform, i.e. the
 * reference's nufht_batch(nu, ws, y, rs) with sources and targets swapped. */
int nufht_adjoint_device(const NufhtPlan *plan,
                         const double *d_ys, int ld_ys,
                         double *d_xs, int ld_xs,
                         int batch, void *stream);
int nufht_adjoint(const NufhtPlan *plan, const double *ys, int ld_ys,
                  double *xs, int ld_xs, int batch);

/* ---- node gradients ---------------------------------------------------------
 * Given rho = dL/dg (n x batch) and the coefficients cs (m x batch):
 *   grad_r[j] = sum_t cs[j,t] sum_i rho[i,t] ws_i J'_nu(ws_i rs_j)
 *   grad_w[i] = sum_t rho[i,t] sum_j cs[j,t] rs_j J'_nu(ws_i rs_j)
 * using J'_nu(x) = (nu/x) J_nu(x) - J_{nu+1}(x) (DLMF 10.6.2), so each is
 * one extra fast transform of order |nu|+1.  Either output may be NULL. */
int nufht_grad_nodes_device(const NufhtPlan *plan,
                            const double *d_cs, int ld_cs,
                            const double *d_rho, int ld_rho, int batch,
                            double *d_grad_r, double *d_grad_w, void *stream);

/* ---- uneven multi-transform batch -------------------------------------------
 * Many independent transforms (own nu, tol, grids, batch) executed with one
 * kernel launch per stage over a shared work queue; optionally replayed as
 * a captured hipGraph.  Plans must outlive the NufhtMulti. */
typedef struct NufhtMulti NufhtMulti;
int  nufht_multi_create(NufhtMulti **out, int count,
                        const NufhtPlan *const *plans, const int *batch);
/* d_cs[k] (ld_cs[k] >= m_k) and d_gs[k] (ld_gs[k] >= n_k) per transform.
 * use_graph != 0 captures the launch sequence once per distinct pointer set
 * and replays it. */
int  nufht_multi_execute(NufhtMulti *mt,
                         const double *const *d_cs, const int *ld_cs,
                         double *const *d_gs, const int *ld_gs,
                         void *stream, int use_graph);
void nufht_multi_free(NufhtMulti *mt);

/* ---- per-stage timing ---------------------------------------------------------
 * When enabled, every stage launch is bracketed by HIP events on its stream;
 * nufht_amd_profile_read returns accumulated milliseconds and launch counts
 * per stage (after synchronising), and the stage names. */
int  nufht_amd_profile(int enable);
int  nufht_amd_profile_read(double *ms, long long *launches, int max_stages,
                            const char **names);
void nufht_amd_profile_reset(void);

/* ---- plan statistics (roofline accounting) ------------------------------------- */
typedef struct {
    long long n_loc, n_asy, n_dir;          /* execution boxes */
    long long pairs_dir;                     /* sum over direct boxes of ni*nj */
    long long loc_src, loc_tgt;              /* sum of nj, ni over local boxes */
    long long asy_src, asy_tgt;              /* sum of nj, ni over asymptotic boxes */
    long long asy_grid;                      /* sum of FFT lengths per channel set */
    long long asy_grid_points;               /* sum of E*B*Nf (complex points) */
    long long asy_channels;                  /* sum of E*B */
    int kernel_width;                        /* ES kernel width (points) */
    int precision;                           /* 0 FP64, 2 mixed */
    int K_asy, K_loc;
    double z_split;
    long long fft_classes;                   /* distinct FFT lengths */
    /* (source, channel) pairs spread and (target, channel) pairs interpolated
     * by the VALU and the matrix-core kernels; a channel is one complex
     * (exponent, column) series, each pair costs kernel_width complex
     * multiply-adds */
    long long spread_sc_valu, spread_sc_mfma;
    long long interp_tc_valu, interp_tc_mfma;
} NufhtPlanStats;
int nufht_amd_plan_stats(const NufhtPlan *plan, int batch, NufhtPlanStats *st);

/* ---- autotuner --------------------------------------------------------------------
 * Times candidate tile shapes / grid oversampling on the device with the
 * plan's own geometry and keeps the fastest for this plan (and batch). */
int nufht_amd_autotune(NufhtPlan *plan, int batch, void *stream);

/* ---- diagnostics ---------------------------------------------------------------- */
const char *nufht_amd_version(void);
const char *nufht_amd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* NUFHT_AMD_H */
