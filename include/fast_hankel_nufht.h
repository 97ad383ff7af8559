/**
 * \file fast_hankel_nufht.h
 * \brief Drop-in C ABI of the batched nonuniform fast Hankel transform,
 *        executed on an AMD Instinct MI355X (gfx950) through HIP.
 *
 *   g[i,t] = sum_j cs[j,t] * J_nu(ws[i] * rs[j]),  i < n, j < m, t < batch.
 *
 * Source-compatible with the reference header (Ansal06/NUFHT
 * src/fast_hankel_nufht.h:14-172): the same type names, fields, function
 * names, argument meanings, column-major layouts with leading dimensions,
 * and negative error codes.  Callers recompile; nothing else changes.
 *
 * Differences a caller can observe (all documented in INTEGRATION.md):
 *  - NufhtPlan and NufhtScratch each append one trailing member, `impl`,
 *    holding device state.  Both remain complete types, so stack
 *    allocation as in the reference's callers keeps working.
 *  - Inputs no longer need to be sorted: the plan sorts rs/ws internally
 *    and permutes cs rows and gs rows back.  Sorted input produces exactly
 *    the reference's box partition (same indices).
 *  - Failures of the device path return -500..-599 instead of printing and
 *    returning 0 (the reference swallows FINUFFT errors,
 *    src/expansions.c:515-518).
 */
#ifndef FAST_HANKEL_NUFHT_H
#define FAST_HANKEL_NUFHT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------ Options (reference src/fast_hankel_nufht.h:21-28) ---- */
typedef struct {
    double tol;          /* default 1e-8 */
    int    max_levels;   /* <= 0 => auto */
    int    min_dim_prod; /* default 10000 */
    double z_split;      /* NaN => table */
    int    K_asy;        /* <= 0 => table */
    int    K_loc;        /* <= 0 => table */
} NufhtOptions;

/** Defaults; replaces reference src/fast_hankel_nufht.c:748-757. */
NufhtOptions nufht_default_options(void);

/* ------------------ Box types (reference h:36-53) ------------------------ */
typedef struct {
    int i0, i1;   /* target (w) index range, inclusive, 0-based */
    int j0, j1;   /* source (r) index range, inclusive, 0-based */
} Box;

typedef struct {
    Box   *data;
    size_t length;
    size_t capacity;
} BoxVector;

typedef struct {
    BoxVector loc;  /* local (Wimp/Chebyshev) expansion boxes */
    BoxVector asy;  /* asymptotic (Hankel) expansion boxes */
    BoxVector dir;  /* direct-summation boxes */
} Boxes;

/** Replaces reference src/fast_hankel_nufht.c:442-450. */
void boxes_free(Boxes *boxes);

/* ------------------ Plan (reference h:62-81) ----------------------------- */
typedef struct {
    double nu;
    double tol;
    double z_split;
    int    K_asy;
    int    K_loc;
    double asy_coef[64];  /* Hankel coefficients a_k(|nu|), k = 0..21 */

    int m;                /* number of sources (rs, cs rows) */
    int n;                /* number of targets (ws, gs rows) */
    const double *rs;     /* not owned; must outlive the plan */
    const double *ws;     /* not owned; must outlive the plan */

    int max_levels;
    int min_dim_prod;

    Boxes boxes;          /* partition of the sorted index rectangle */

    int tables_loaded;

    void *impl;           /* appended: sorted grids, device program */
} NufhtPlan;

/* ------------------ Scratch (reference h:84-97) -------------------------- */
typedef struct {
    double *in_buffer;
    double *out_buffer;
    double *real_buffer_1;
    double *real_buffer_2;
    double *cheb_buffer;
    double *bessel_buffer_1;
    double *bessel_buffer_2;

    int m_alloc;
    int n_alloc;
    int Kloc_alloc;
    int extra_alloc;

    void *impl;           /* appended: device staging for host-pointer calls */
} NufhtScratch;

/** Replaces reference src/fast_hankel_nufht.c:759-786.
 *  Returns 0, -200 (bad grids), -1 (tol < 1e-15), -2 (unsupported nu: not
 *  an integer |nu| <= 200 or half-integer |nu| <= 19/2, or a negative
 *  half-integer, which the reference answers with J_{|nu|}), -3..-6 (table
 *  index), -100 (NULL options path), -501 (out of memory). */
int nufht_plan_init(NufhtPlan *plan,
                    double nu,
                    const double *rs, int m,
                    const double *ws, int n,
                    const NufhtOptions *opt);

/** Replaces reference src/fast_hankel_nufht.c:788-793. */
void nufht_plan_free(NufhtPlan *plan);

/** Replaces reference src/fast_hankel_nufht.c:802-832 (-300, -301). */
int nufht_scratch_init(NufhtScratch *scratch, const NufhtPlan *plan);
/** Replaces reference src/fast_hankel_nufht.c:834-845. */
void nufht_scratch_free(NufhtScratch *scratch);

/* ------------------ Box generation (reference h:123-137) ----------------- */
/** Replaces reference src/fast_hankel_nufht.c:578-744 (same partition on
 *  sorted input; boundary searches are binary instead of linear). */
Boxes generate_boxes(const double *rs, int m,
                     const double *ws, int n,
                     double z_split,
                     int max_levels,
                     int min_dim_prod,
                     int K_loc_threshold);

/** Replaces reference src/fast_hankel_nufht.c:523-574. */
void split_box(const double *rs, int m,
               const double *ws, int n,
               const Box    *box,
               double        z,
               int          *ispl_out,
               int          *jspl_out);

/* ------------------ Execution (reference h:148-168) ---------------------- */
/** Replaces reference src/fast_hankel_nufht.c:989-1107.  Host pointers;
 *  cs is batch columns of ld_cs >= m doubles, gs batch columns of
 *  ld_gs >= n doubles, fully overwritten.  scratch may be NULL.
 *  Returns 0, -400 (NULL / batch <= 0), -401 (ld too small),
 *  -402/-403 (allocation), or -5xx (device failure). */
int nufht_batch(const NufhtPlan *plan,
                const double *cs, int ld_cs,
                double *gs, int ld_gs,
                int batch,
                NufhtScratch *scratch);

/** Replaces reference src/fast_hankel_nufht.c:1111-1144. */
void nufht_inplace(double       *gs,
                   double        nu,
                   const double *rs, int m,
                   const double *cs,
                   const double *ws, int n,
                   const NufhtOptions *opt);

/** Replaces reference src/fast_hankel_nufht.c:1148-1159 (caller frees). */
double *nufht(double        nu,
              const double *rs, int m,
              const double *cs,
              const double *ws, int n,
              const NufhtOptions *opt);

#ifdef __cplusplus
}
#endif

#endif /* FAST_HANKEL_NUFHT_H */
