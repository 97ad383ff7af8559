/*
 * Internal interface between the C host planner (plan.c, tables.c) and the
 * HIP device engine (engine.hip).  Not installed; not part of the ABI.
 */
#ifndef NUFHT_AMD_INTERNAL_H
#define NUFHT_AMD_INTERNAL_H

#include "fast_hankel_nufht.h"
#include "nufht_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host half of NufhtPlan.impl. */
typedef struct NufhtImplHost {
    double *r, *w;          /* ascending copies of rs / ws (owned) */
    int *perm_r, *perm_w;   /* sorted position -> caller index; NULL if the
                               caller's grid was already ascending */
    NufhtOptions opt;       /* options the plan was built with */
    int precision;          /* NUFHT_PREC_* requested */
    /* execution partition: equals plan->boxes for integer orders; for
     * half-integer orders it is built with exec_z (see plan.c) */
    Boxes exec_boxes;
    double exec_z;          /* split used by exec_boxes */
    int exec_K_loc;         /* local expansion terms used by exec_boxes */
    int exec_K_asy;         /* asymptotic terms */
    int use_direct;         /* m*n < min_dim_prod: one direct box */
    void *dev;              /* engine state (engine.hip) */
    struct NufhtPlanDerived {
        NufhtPlan *adjoint;      /* (|nu|, ws, rs) */
        NufhtPlan *adjoint_up;   /* (|nu|+1, ws, rs) */
        NufhtPlan *forward_up;   /* (|nu|+1, rs, ws) */
    } derived;
} NufhtImplHost;

/* tables.c */
int  nufht_tables_row(double abs_nu, double *coef64);
int  nufht_tables_lookup(int nu_int, int j_idx, int k_slot, double *z, int *wimp_k);
double nufht_asy_error_bound(double nu, int K, double z);

/* plan.c helpers used by the engine */
/* Largest uniform grid (FFT length) of one asymptotic box: 2^30 points. */
#define NUFHT_ASY_MAX_NF (1LL << 30)
/* Grid of an asymptotic box with source half-width X and grid spacing h for
 * an ES kernel of width kw: half-width L (positions -L..L) and FFT length
 * Nf.  Returns 0, or -541 (with the error text set) when Nf would exceed
 * NUFHT_ASY_MAX_NF. */
int nufht_asy_grid(double X, double h, int kw, int *L, int *Nf);
int nufht_build_partition(const double *r, int m, const double *w, int n,
                          double z, int max_levels, int min_dim_prod, int kloc,
                          Boxes *out);

/* scratch.hip: the allocation-free entry point's static device array (its
 * address on the current device and its length in doubles) */
int  nufht_scratch_array(double **out, long long *count);

/* engine.hip */
void nufht_engine_release(NufhtImplHost *h);
int  nufht_engine_batch_host(const NufhtPlan *plan, const double *cs, int ld_cs,
                             double *gs, int ld_gs, int batch, NufhtScratch *scratch);
void nufht_engine_scratch_release(NufhtScratch *scratch);
void nufht_set_error(const char *fmt, ...);

#ifdef __cplusplus
}
#endif
#endif
