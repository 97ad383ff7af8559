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
 * AUTO (default): when tol >= 1e-6, the asymptotic stage's grids are FP32:
 *   spreading multiplies and accumulates in FP32, the uniform FFT runs in
 *   FP32 and interpolation reads FP32 grid values; phases, deconvolution,
 *   the interpolation's products and sums, Bessel evaluation, the local
 *   expansions and the direct sums stay FP64.  tol < 1e-6 runs FP64.
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

/* ---- allocation-free execution ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory during a call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the transform of nufht_batch_device,
 * calling no hipMalloc, hipFree or allocating library routine.  The kernels'
 * scratch (local-expansion moments; the sorted-order copies of cs and gs
 * when the plan's grids were not ascending) lives in a static device array
 * of 2^24 doubles (128 MiB) that each GPU the library runs on holds until
 * the process ends.  The plan's launch program must exist already: run
 * nufht_batch_device once with the same batch (and the plan's current
 * precision and knobs) first; that call allocates the plan's own state.
 * All calls on one GPU share that GPU's scratch array: calls that overlap
 * on one GPU (other streams, other threads, other plans) race; order them.
 * Returns 0, -400/-401 (as nufht_batch_device), -550 (the scratch this call
 * needs, moments plus m*batch and/or n*batch staged values, exceeds 2^24
 * doubles), -551 (no launch program for this batch yet), or -5xx. */
int nufht_batch_device_noalloc(const NufhtPlan *plan,
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
/* Where calls fork the near-field stages (local expansions, direct sums)
 * onto the program's second stream: 1 before the spreads (default), 2 before
 * the FFTs, 0 never (one stream: every stage then runs alone, so per-stage
 * timings measure each kernel without the other stream's contention).  Takes
 * effect for calls made afterwards; at < 0 only queries.  Returns the
 * previous setting.  (NUFHT_AMD_FORK sets the initial value.) */
int  nufht_amd_stream_fork(int at);
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
    /* the roofline model's fine grid per asymptotic box, nf_b = 2 sigma X S / pi
     * + kernel_width + 1 (sigma = 2; X, S the box's r and k half-widths),
     * weighted by the box's channel count C_b: sum C_b nf_b and
     * sum C_b nf_b log2 nf_b */
    double asy_fine, asy_fine_nlogn;
    /* bytes the uniform FFTs move per call: for the two-pass kernels, pass 1
     * reads a channel's 2L + 1 spread positions and writes its Nf points,
     * pass 2 reads Nf and writes the frequencies interpolation reads; for
     * rocFFT classes one read and one write of Nf points */
    double fft_bytes;
    /* the roofline model's algorithmic work of one call: FP64 flops, FP32
     * flops, HBM bytes.  Direct sums 2 batch flops per (target, source) pair
     * (Bessel evaluation not counted); local expansions 2 (K_loc + 1) batch
     * per source and per target; per asymptotic box with C channels,
     * 2 w C per source (spread), 4 w C per target (interpolation) and
     * 2.5 C nf_b log2 nf_b (FFT), FP32 in mixed precision; bytes 8 (m + n)
     * (1 + batch) plus 16 C nf_b per box (8 in mixed precision).  The
     * floor of one call on MI355X is max(f64 / 78.6e12 + f32 / 157.3e12,
     * bytes / 8e12) seconds. */
    double model_f64, model_f32, model_bytes;
    /* the direct-sum and local-expansion shares of model_f64 */
    double model_dir, model_loc;
    /* asymptotic boxes run by the fused small-box kernel, and their
     * (source, channel) and (target, channel) pairs (not in the spread_sc_*
     * and interp_tc_* counts) */
    long long fused_boxes, fused_sc, fused_tc;
    /* local-expansion evaluation items that loop over every box covering
     * their targets (one atomic add per output element per item) */
    long long loc_eval_multi;
} NufhtPlanStats;
int nufht_amd_plan_stats(const NufhtPlan *plan, int batch, NufhtPlanStats *st);
/* The same statistics summed over the transforms of a multi-transform batch
 * (kernel_width, K_asy, K_loc and z_split are the first transform's). */
int nufht_multi_stats(const NufhtMulti *mt, NufhtPlanStats *st);

/* ---- launch-geometry knobs and the autotuner -----------------------------------
 * Choices that change which kernels run and how work is cut into items, never
 * the arithmetic's accuracy target:
 *   dense_cell    sources per grid position from which an asymptotic box
 *                 spreads on the matrix cores (else the VALU spread kernel)
 *   dense_split   consecutive sources per matrix-core spread item (0: 512 for
 *                 kernel widths <= 8, else 1024)
 *   spread_split  sources per VALU spread item
 *   interp_mfma   1: 16-target tiles whose kernel supports fit a 16-position
 *                 window interpolate on the matrix cores, in runs of up to 8
 *                 tiles whose window is staged in LDS once; 2: the same tiles
 *                 one per item, windows read from HBM; 0: VALU only
 *   own_fft       1: two-pass FFT kernels for the lengths they cover; 0: rocFFT
 *   spread_sched  matrix-core spread schedule: 1 split (4 waves stage source
 *                 chunks into double-buffered LDS while 4 multiply), 0 unified
 *                 (4 waves stage, then multiply), -1 automatic (split)
 *   fused_max     asymptotic boxes with a uniform grid of at most 256 points
 *                 and at most fused_max sources and targets run the fused
 *                 kernel: spread, FFT and interpolation with the grid in LDS,
 *                 in one persistent launch over a work queue of their units
 *                 (default 0: every box runs the staged kernels, measured
 *                 faster on the benchmark workloads; NUFHT_AMD_FUSED_MAX
 *                 overrides the default)
 *   precision     one of NUFHT_PREC_* (as nufht_amd_plan_set_precision)
 * A plan starts from the defaults (NUFHT_AMD_DENSE_CELL and
 * NUFHT_AMD_PRECISION override two of them). */
typedef struct {
    double dense_cell;
    int dense_split, spread_split;
    int interp_mfma, own_fft;
    int precision;
    int spread_sched;
    int fused_max;
} NufhtAmdKnobs;
int nufht_amd_get_knobs(const NufhtPlan *plan, NufhtAmdKnobs *k);
int nufht_amd_set_knobs(NufhtPlan *plan, const NufhtAmdKnobs *k);

/* Times the knobs on the device with the plan's own geometry and batch
 * (coordinate descent: precision when tol >= 1e-6, then dense_cell,
 * interp_mfma, dense_split, spread_split, own_fft, spread_sched, fused_max; each candidate is built,
 * run once untimed and timed over three calls) and keeps the fastest. */
int nufht_amd_autotune(NufhtPlan *plan, int batch, void *stream);
/* The same for a multi-transform batch (precision stays per plan); the
 * launch sequence is rebuilt and, with graphs, captured again. */
int nufht_multi_autotune(NufhtMulti *mt, void *stream);
int nufht_multi_get_knobs(const NufhtMulti *mt, NufhtAmdKnobs *k);

/* ---- table cache writer ------------------------------------------------------------
 * Replaces the reference's table generation and cache save
 * (src/bounds.c:306-324, src/fast_hankel_nufht.c:193-247): computes every
 * entry the reference's generator fills and writes them in its cache format
 * (the file NUFHT_TABLE_CACHE may name, and the reference's
 * data/nufht_tables.cache).  Returns 0, -1 (memory) or -2 (I/O). */
int nufht_amd_write_table_cache(const char *path);

/* ---- plan cache -------------------------------------------------------------------
 * No counterpart in the reference, which plans on every call.  The device
 * state of a freed plan (sorted nodes, launch program, grids, FFT plans,
 * tuned knobs) is kept for reuse by the next plan with the same order,
 * options, sizes and node arrays (a 128-bit hash of r and of k), so that
 * callers that plan every time (the _pynufht entry points) pay the device
 * set-up once.  Keeps the most recent `capacity` states when capacity >= 0
 * (default 2, NUFHT_AMD_PLAN_CACHE; 0 frees them and disables the cache);
 * *hits (may be NULL) receives the number of reuses so far. */
int nufht_amd_plan_cache(int capacity, long long *hits);

/* ---- diagnostics ---------------------------------------------------------------- */
const char *nufht_amd_version(void);
const char *nufht_amd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* NUFHT_AMD_H */
