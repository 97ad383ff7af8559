/*
 * oracle/oracle.h -- TEST INFRASTRUCTURE (the oracle), never product code.
 *
 * A plain-C, CPU-only restatement of the reference's nufht_batch path
 * (Ansal06/NUFHT src/fast_hankel_nufht.c, src/expansions.c, src/bounds.c)
 * plus an FP64 direct-sum reference.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only to check results or
 * as the timed CPU baseline.  Nothing in nufht_amd/ includes or links it.
 */
#ifndef NUFHT_ORACLE_H
#define NUFHT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Bessel (bessel.c) ---- */
double orc_jn_int_dd(int n, double hi, double lo);
double orc_jn_half_dd(int n, double hi, double lo);
double orc_besselj_wr(double nu, double w, double r);
double orc_besselj(double nu, double x);

/* ---- FP64 direct sum (direct.c) ----
 * g[i + t*ldg] = sum_j c[j + t*ldc] * J_nu(w_i r_j), Neumaier-compensated,
 * OpenMP over targets.  nu follows the reference's sign conventions
 * (negative integer: (-1)^nu J_|nu|; negative half-integer: J_|nu|). */
void orc_direct(double nu, const double *r, int m, const double *c, int ldc,
                int B, const double *w, int n, double *g, int ldg,
                int nthreads);
/* Same on a subset of target rows rows[0..nr-1]; out is nr x B column-major. */
void orc_direct_rows(double nu, const double *r, int m, const double *c,
                     int ldc, int B, const double *w, const int *rows, int nr,
                     double *out, int nthreads);

/* ---- plan tables (tables.c), restating src/bounds.c:63-324 ---- */
#define ORC_NU_ROWS   401
#define ORC_ASY_COEF  64
#define ORC_TI        401
#define ORC_TJ        32
#define ORC_TK        16
double orc_a_coef(int k, double nu);
double orc_asy_error_bound(double nu, int K, double z);
double orc_wimp_error_bound(double nu, int K, double z);
/* Fills the three tables exactly as the reference's generate_tables does
 * (quietly).  Arrays are [401][64], [401][32][16], [401][32][16]. */
void orc_generate_tables(double *asy_coef, double *asy_z, int *wimp_k);

/* ---- plan + boxes (plan.c), restating src/fast_hankel_nufht.c ---- */
typedef struct { int i0, i1, j0, j1; } OrcBox;
typedef struct {
    OrcBox *loc, *asy, *dir;
    int nloc, nasy, ndir;
} OrcBoxes;
typedef struct {
    double nu, tol, z_split;
    int K_asy, K_loc;
    double asy_coef[64];
    int max_levels, min_dim_prod;
} OrcParams;
int orc_fill_params(OrcParams *p, double nu, double tol, int max_levels,
                    int min_dim_prod, double z_split, int K_asy, int K_loc);
int orc_generate_boxes(const double *rs, int m, const double *ws, int n,
                       double z_split, int max_levels, int min_dim_prod,
                       int K_loc_threshold, OrcBoxes *out);
void orc_boxes_free(OrcBoxes *b);
void orc_split_box(const double *rs, int m, const double *ws, int n,
                   const OrcBox *box, double z, int *ispl, int *jspl);

/* ---- ref-algorithm CPU port (port.c) ----
 * nufht_batch with the reference's plan, boxes and per-regime algebra.
 * asy_mode 0: exact exponential sums in the asymptotic regime (pins the
 * algebra, O(n_i n_j)); 1: the oracle's own CPU type-3 NUFFT at eps = tol
 * (the timed CPU baseline).  Returns the reference's error codes. */
/* planted mistakes of the port (tests only; 0 = none, see port.c) */
void orc_set_plant(int which);
int orc_nufht_batch(double nu, const double *rs, int m, const double *ws,
                    int n, const double *cs, int ldc, double *gs, int ldg,
                    int B, double tol, int asy_mode, int nthreads);

/* ---- CPU type-3 NUFFT (nufft.c) ----
 * f[k + t*n] = sum_j c[j + t*m] exp(i*iflag*x_j*s_k), complex interleaved. */
int orc_nufft3(int m, const double *x, const double *c, int iflag, double eps,
               int n, const double *s, double *f, int ntr);
void orc_nufft3_direct(int m, const double *x, const double *c, int iflag,
                       int n, const double *s, double *f, int ntr);

#ifdef __cplusplus
}
#endif
#endif
