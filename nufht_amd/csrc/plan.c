/*
 * Host plan builder (C): options, plan parameters, the (w, r) box
 * partition, sorting, and the plan/scratch lifecycle of the drop-in ABI.
 *
 *   nufht_default_options  reference src/fast_hankel_nufht.c:748-757
 *   plan parameters        reference src/fast_hankel_nufht.c:322-414
 *   generate_boxes         reference src/fast_hankel_nufht.c:578-744
 *   split_box              reference src/fast_hankel_nufht.c:498-574
 *   plan / scratch         reference src/fast_hankel_nufht.c:759-845
 *   nufht_batch            reference src/fast_hankel_nufht.c:989-1107
 *   nufht_inplace / nufht  reference src/fast_hankel_nufht.c:1111-1159
 *
 * The partition rule is the reference's; the boundary searches ("first
 * index whose value exceeds t") are binary searches over the ascending
 * grids, which return the same index as the reference's linear scans.
 * Grids are sorted (with the permutation kept) before planning, so callers
 * may pass unsorted nodes.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

NufhtOptions nufht_default_options(void)
{
    NufhtOptions o;
    o.tol = 1e-8;
    o.max_levels = 0;
    o.min_dim_prod = 10000;
    o.z_split = NAN;
    o.K_asy = 0;
    o.K_loc = 0;
    return o;
}

static int is_int(double x) { return fabs(x - nearbyint(x)) < 1e-12; }

static int ascending(const double *v, int n)
{
    for (int i = 1; i < n; ++i)
        if (!(v[i - 1] <= v[i])) return 0;
    return 1;
}

/* first index i with v[i] > t, or -1 (reference find_first_*_greater) */
static int first_greater_sorted(const double *v, int n, double t)
{
    if (isnan(t)) return -1;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (v[mid] > t) hi = mid;
        else lo = mid + 1;
    }
    return lo < n ? lo : -1;
}

static int first_greater_scan(const double *v, int n, double t)
{
    for (int i = 0; i < n; ++i)
        if (v[i] > t) return i;
    return -1;
}

typedef int (*finder)(const double *, int, double);

static void bv_push(BoxVector *v, int i0, int i1, int j0, int j1)
{
    if (v->length == v->capacity) {
        size_t cap = v->capacity ? 2 * v->capacity : 8;
        Box *d = (Box *)realloc(v->data, cap * sizeof(Box));
        if (!d) return;
        v->data = d;
        v->capacity = cap;
    }
    Box b = {i0, i1, j0, j1};
    v->data[v->length++] = b;
}

void boxes_free(Boxes *b)
{
    if (!b) return;
    free(b->loc.data);
    free(b->asy.data);
    free(b->dir.data);
    memset(b, 0, sizeof(*b));
}

static void split_with(finder ff, const double *rs, int m, const double *ws,
                       const Box *box, double z, int *ispl, int *jspl)
{
    int cand[10], nc = 0;
    if (box->i1 <= box->i0) {
        cand[nc++] = box->i0;
    } else {
        for (int k = 0; k < 10; ++k) {
            int idx = (int)llround(box->i0 + ((double)k / 9.0) * (double)(box->i1 - box->i0));
            if (idx < box->i0) idx = box->i0;
            if (idx > box->i1) idx = box->i1;
            if (nc == 0 || idx != cand[nc - 1]) cand[nc++] = idx;
        }
    }
    long long best = -1;
    int bi = box->i0;
    for (int c = 0; c < nc; ++c) {
        int i = cand[c];
        int j = ff(rs, m, z / ws[i]);
        long long v = 0;
        if (j >= 0)
            v = (long long)(i - box->i0 + 1) * (long long)(j - box->j0 + 1) +
                (long long)(box->i1 - i + 1) * (long long)(box->j1 - j + 1);
        if (v > best) { best = v; bi = i; }
    }
    int j = ff(rs, m, z / ws[bi]);
    *ispl = bi;
    *jspl = j < 0 ? box->j0 : j;
}

void split_box(const double *rs, int m, const double *ws, int n, const Box *box,
               double z, int *ispl, int *jspl)
{
    (void)n;
    split_with(ascending(rs, m) ? first_greater_sorted : first_greater_scan, rs, m, ws,
               box, z, ispl, jspl);
}

static Boxes partition(finder fr, finder fw, const double *rs, int m, const double *ws,
                       int n, double z, int max_levels, int min_dim_prod, int kloc)
{
    Boxes bx;
    memset(&bx, 0, sizeof(bx));
    if (n <= 0 || m <= 0) return bx;
    if (max_levels <= 0) {
        int mn = n < m ? n : m;
        double ratio = (double)mn * (double)mn / (double)min_dim_prod;
        max_levels = ratio < 1.0 ? 0 : (int)floor(log2(ratio));
        if (max_levels < 0) max_levels = 0;
    }
    int i0 = fw(ws, n, z / rs[m - 1]), i1, j0, j1;
    if (m == 1) {
        i1 = i0 == -1 ? -1 : i0 - 1;
    } else {
        int idx = fw(ws, n, z / rs[0]);
        i1 = idx == -1 ? n - 1 : idx - 1;
    }
    j0 = fr(rs, m, z / ws[n - 1]);
    if (n == 1) {
        j1 = j0 == -1 ? -1 : j0 - 1;
    } else {
        int idx = fr(rs, m, z / ws[0]);
        j1 = idx == -1 ? m - 1 : idx - 1;
    }
    if (i0 == -1) {
        bv_push(&bx.loc, 0, n - 1, 0, m - 1);
        return bx;
    }
    if (i1 < 0) {
        bv_push(&bx.asy, 0, n - 1, 0, m - 1);
        return bx;
    }
    if (i0 > 0) bv_push(&bx.loc, 0, i0 - 1, 0, m - 1);
    if (j0 > 0) bv_push(&bx.loc, i0, n - 1, 0, j0 - 1);
    if (i1 < n - 1) bv_push(&bx.asy, i1 + 1, n - 1, 0, m - 1);
    if (j1 < m - 1) bv_push(&bx.asy, 0, i1, j1 + 1, m - 1);
    if (m == 1 || n == 1) return bx;
    bv_push(&bx.dir, i0, i1, j0, j1);
    for (int level = 0; level < max_levels; ++level) {
        BoxVector nd = {0};
        for (size_t q = 0; q < bx.dir.length; ++q) {
            Box b = bx.dir.data[q];
            long long d1 = b.i1 - b.i0 + 1, d2 = b.j1 - b.j0 + 1;
            if (d1 * d2 > 4LL * (long long)min_dim_prod) {
                int is, js;
                split_with(fr, rs, m, ws, &b, z, &is, &js);
                if (is > b.i0) {
                    bv_push(&bx.loc, b.i0, is - 1, b.j0, js - 1);
                    bv_push(&nd, b.i0, is - 1, js, b.j1);
                }
                bv_push(&bx.asy, is, b.i1, js, b.j1);
                bv_push(&nd, is, b.i1, b.j0, js - 1);
            } else {
                bv_push(&nd, b.i0, b.i1, b.j0, b.j1);
            }
        }
        free(bx.dir.data);
        bx.dir = nd;
    }
    BoxVector nl = {0}, na = {0};
    for (size_t q = 0; q < bx.loc.length; ++q) {
        Box b = bx.loc.data[q];
        int d1 = b.i1 - b.i0 + 1, d2 = b.j1 - b.j0 + 1;
        if ((d1 < d2 ? d1 : d2) < kloc) bv_push(&bx.dir, b.i0, b.i1, b.j0, b.j1);
        else bv_push(&nl, b.i0, b.i1, b.j0, b.j1);
    }
    for (size_t q = 0; q < bx.asy.length; ++q) {
        Box b = bx.asy.data[q];
        long long area = (long long)(b.i1 - b.i0 + 1) * (long long)(b.j1 - b.j0 + 1);
        if (area < (long long)min_dim_prod) bv_push(&bx.dir, b.i0, b.i1, b.j0, b.j1);
        else bv_push(&na, b.i0, b.i1, b.j0, b.j1);
    }
    free(bx.loc.data);
    free(bx.asy.data);
    bx.loc = nl;
    bx.asy = na;
    return bx;
}

Boxes generate_boxes(const double *rs, int m, const double *ws, int n, double z_split,
                     int max_levels, int min_dim_prod, int K_loc_threshold)
{
    finder fr = ascending(rs, m) ? first_greater_sorted : first_greater_scan;
    finder fw = ascending(ws, n) ? first_greater_sorted : first_greater_scan;
    return partition(fr, fw, rs, m, ws, n, z_split, max_levels, min_dim_prod,
                     K_loc_threshold);
}

int nufht_build_partition(const double *r, int m, const double *w, int n, double z,
                          int max_levels, int min_dim_prod, int kloc, Boxes *out)
{
    *out = partition(first_greater_sorted, first_greater_sorted, r, m, w, n, z, max_levels,
                     min_dim_prod, kloc);
    return 0;
}

/* FFT length for a grid of `need` points: the smallest 2^a or 3 2^a, or
 * above 4096 (the two-pass FFT kernels' range) 9 2^a, that holds it. */
static long long fft_length_for(long long need)
{
    long long best = -1;
    for (long long p = 8; p <= NUFHT_ASY_MAX_NF; p <<= 1) {
        if (p >= need) {
            if (best < 0 || p < best) best = p;
            break;
        }
        if (3 * p >= need && (best < 0 || 3 * p < best)) best = 3 * p;
        if (9 * p >= need && 9 * p > 4096 && (best < 0 || 9 * p < best)) best = 9 * p;
    }
    return best;
}

int nufht_asy_grid(double X, double h, int kw, int *L, int *Nf)
{
    /* type-3 NUFFT grid of an asymptotic box (upsampling sigma = 2):
     * positions l in [-L, L], L = ceil(X / h) + kw / 2 + 2, FFT length
     * Nf >= sigma (2 L + 1) and >= 2 kw + 8 */
    const double sigma = 2.0;
    double Lf = ceil(X / h) + (double)(kw / 2) + 2.0;
    if (!(Lf >= 0.0) || !(sigma * (2.0 * Lf + 1.0) <= (double)NUFHT_ASY_MAX_NF)) {
        nufht_set_error("an asymptotic box needs a uniform grid of %.3g points (half-width "
                        "X = %.6g, spacing h = %.6g): above the supported %lld",
                        sigma * (2.0 * Lf + 1.0), X, h, (long long)NUFHT_ASY_MAX_NF);
        return -541;
    }
    long long l = (long long)Lf, need = (long long)(sigma * (double)(2 * l + 1));
    if (need < 2LL * kw + 8) need = 2LL * kw + 8;
    long long nf = fft_length_for(need);
    if (nf < 0) {
        nufht_set_error("no FFT length of at most %lld holds %lld points",
                        (long long)NUFHT_ASY_MAX_NF, need);
        return -541;
    }
    *L = (int)l;
    *Nf = (int)nf;
    return 0;
}

/* reference fill_plan_parameters, src/fast_hankel_nufht.c:322-414 */
static int fill_params(NufhtPlan *p, double nu, const NufhtOptions *o)
{
    if (o->tol < 1e-15) {
        nufht_set_error("tolerance below 1e-15");
        return -1;
    }
    int two_int = is_int(2.0 * nu), nu_int = is_int(nu);
    if (!two_int || (nu_int && fabs(nu) > 200.0) || (!nu_int && fabs(nu) > 9.5)) {
        nufht_set_error("only integer |nu| <= 200 and half-integer |nu| <= 19/2 are supported");
        return -2;
    }
    /* The reference accepts nu = -(n + 1/2) and computes J_{n+1/2} (it takes
     * |nu| everywhere, src/fast_hankel_nufht.c:336-344, :932).  J_{-(n+1/2)}
     * is a different function (DLMF 10.47.3 with y_n, singular at 0), so
     * such orders are refused instead of answered with the wrong one. */
    if (!nu_int && nu < 0.0) {
        nufht_set_error("negative half-integer orders are not supported: J_{-(n+1/2)} is "
                        "singular at 0 and is not J_{n+1/2}");
        return -2;
    }
    int i_idx = (int)llround(fabs(nu));
    if (i_idx < 0 || i_idx >= 401) return -3;
    int row = (int)llround(2.0 * fabs(nu) + 1.0) - 1;
    if (row < 0 || row >= 401) return -4;
    nufht_tables_row(fabs(nu), p->asy_coef);
    int j = (int)ceil(-log10(o->tol) - 3.0);
    if (j < 1) j = 1;
    int j_idx = j - 1;
    if (j_idx < 0 || j_idx >= 32) return -5;
    p->nu = nu;
    p->tol = o->tol;
    if (nu_int) {
        int k;
        if (o->K_asy <= 0) {
            int k0 = (int)floor(fabs(nu) / 5.0 + log10(1.0 / o->tol) / 4.0 + 1.0);
            k = k0 < 10 ? k0 : 10;
        } else {
            k = o->K_asy;
        }
        if (k < 0 || k >= 16) return -6;
        p->K_asy = k;
        double z;
        int wk;
        if (k <= 10) {
            nufht_tables_lookup(i_idx, j_idx, k, &z, &wk);
        } else {
            z = 0.0;   /* the reference's table has nothing beyond slot 10 */
            wk = 0;
        }
        p->z_split = isnan(o->z_split) ? z : o->z_split;
        p->K_loc = o->K_loc <= 0 ? wk : o->K_loc;
    } else {
        p->K_asy = (int)llround(fabs(nu) - 0.5);
        p->z_split = NAN;
        p->K_loc = -1;
    }
    p->tables_loaded = 1;
    return 0;
}

/* Chebyshev terms needed to represent j_n(z x), x in [-1, 1], to eps: the
 * coefficients of an entire function of exponential type z decay like
 * J_k(z); take the first k past z where (z/2)^k / k! < eps/8. */
static int cheb_terms_for(double z, double eps)
{
    double t = 1.0;
    int k = 0;
    while (k < 2000) {
        k++;
        t *= (z / 2.0) / (double)k;
        if (k > z && t < eps / 8.0) break;
    }
    return k;
}

static int sort_copy(const double *src, int n, double **dst, int **perm);

int nufht_plan_init(NufhtPlan *plan, double nu, const double *rs, int m, const double *ws,
                    int n, const NufhtOptions *opt_in)
{
    if (!plan || !rs || !ws || m <= 0 || n <= 0) return -200;
    NufhtOptions opt = opt_in ? *opt_in : nufht_default_options();
    memset(plan, 0, sizeof(*plan));
    plan->rs = rs;
    plan->ws = ws;
    plan->m = m;
    plan->n = n;
    plan->max_levels = opt.max_levels;
    plan->min_dim_prod = opt.min_dim_prod;
    int rc = fill_params(plan, nu, &opt);
    if (rc) return rc;

    NufhtImplHost *h = (NufhtImplHost *)calloc(1, sizeof(NufhtImplHost));
    if (!h) return -501;
    h->opt = opt;
    h->precision = NUFHT_PREC_AUTO;
    {
        const char *env = getenv("NUFHT_AMD_PRECISION");
        if (env && !strcmp(env, "fp64")) h->precision = NUFHT_PREC_FP64;
        if (env && !strcmp(env, "mixed")) h->precision = NUFHT_PREC_MIXED;
    }
    if (sort_copy(rs, m, &h->r, &h->perm_r) || sort_copy(ws, n, &h->w, &h->perm_w)) {
        free(h->r); free(h->w); free(h->perm_r); free(h->perm_w); free(h);
        return -501;
    }
    plan->impl = h;
    plan->boxes = partition(first_greater_sorted, first_greater_sorted, h->r, m, h->w, n,
                            plan->z_split, opt.max_levels, opt.min_dim_prod,
                            plan->K_loc < 0 ? 0 : plan->K_loc);
    h->use_direct = (long long)m * (long long)n < (long long)opt.min_dim_prod;
    if (is_int(nu)) {
        h->exec_boxes = plan->boxes;          /* shared, freed once */
        h->exec_z = plan->z_split;
        h->exec_K_loc = plan->K_loc;
        h->exec_K_asy = plan->K_asy;
    } else {
        /* Half-integer orders: the reference sends the whole grid through the
         * finite Hankel series (src/fast_hankel_nufht.c:1052-1078), whose
         * weights r^(-2l-1/2) make it unstable where w*r is small.  The
         * execution partition therefore uses the split of the next integer
         * order, with the near field on a Chebyshev local expansion of
         * j_n(w r) and the far field on the exact finite series. */
        int nu_up = (int)ceil(fabs(nu));
        int k0 = (int)floor(nu_up / 5.0 + log10(1.0 / opt.tol) / 4.0 + 1.0);
        int k = k0 < 10 ? k0 : 10;
        int j = (int)ceil(-log10(opt.tol) - 3.0);
        if (j < 1) j = 1;
        double z;
        int wk;
        nufht_tables_lookup(nu_up, j - 1, k, &z, &wk);
        if (!isnan(opt.z_split)) z = opt.z_split;
        h->exec_z = z;
        h->exec_K_asy = plan->K_asy;
        h->exec_K_loc = cheb_terms_for(z, opt.tol) / 2 + 1;
        h->exec_boxes = partition(first_greater_sorted, first_greater_sorted, h->r, m, h->w,
                                  n, z, opt.max_levels, opt.min_dim_prod, h->exec_K_loc);
    }
    return 0;
}

static int sort_copy(const double *src, int n, double **dst, int **perm)
{
    *dst = (double *)malloc(sizeof(double) * (size_t)n);
    *perm = NULL;
    if (!*dst) return -1;
    memcpy(*dst, src, sizeof(double) * (size_t)n);
    if (ascending(src, n)) return 0;
    int *p = (int *)malloc(sizeof(int) * (size_t)n);
    int *tmp = (int *)malloc(sizeof(int) * (size_t)n);
    if (!p || !tmp) { free(p); free(tmp); return -1; }
    for (int i = 0; i < n; ++i) p[i] = i;
    /* stable merge sort of indices by value */
    for (int width = 1; width < n; width *= 2) {
        for (int lo = 0; lo < n; lo += 2 * width) {
            int mid = lo + width < n ? lo + width : n;
            int hi = lo + 2 * width < n ? lo + 2 * width : n;
            int a = lo, b = mid, o = lo;
            while (a < mid && b < hi) tmp[o++] = (src[p[b]] < src[p[a]]) ? p[b++] : p[a++];
            while (a < mid) tmp[o++] = p[a++];
            while (b < hi) tmp[o++] = p[b++];
        }
        int *sw = p; p = tmp; tmp = sw;
    }
    free(tmp);
    for (int i = 0; i < n; ++i) (*dst)[i] = src[p[i]];
    *perm = p;
    return 0;
}

void nufht_plan_free(NufhtPlan *plan)
{
    if (!plan) return;
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    if (h) {
        nufht_engine_release(h);
        if (h->exec_boxes.loc.data != plan->boxes.loc.data ||
            h->exec_boxes.asy.data != plan->boxes.asy.data ||
            h->exec_boxes.dir.data != plan->boxes.dir.data)
            boxes_free(&h->exec_boxes);
        NufhtPlan *d[3] = {h->derived.adjoint, h->derived.adjoint_up, h->derived.forward_up};
        for (int i = 0; i < 3; ++i)
            if (d[i]) { nufht_plan_free(d[i]); free(d[i]); }
        free(h->r); free(h->w); free(h->perm_r); free(h->perm_w);
        free(h);
    }
    boxes_free(&plan->boxes);
    memset(plan, 0, sizeof(*plan));
}

int nufht_scratch_init(NufhtScratch *s, const NufhtPlan *plan)
{
    if (!s || !plan) return -300;
    memset(s, 0, sizeof(*s));
    int kl = plan->K_loc < 0 ? 0 : plan->K_loc;
    long ni = (long)llround(plan->nu);
    int extra = 0;
    if (plan->K_loc >= 0) {
        extra = (int)(labs(ni) / 2);
        if (ni & 1L) extra += 1;
    }
    s->in_buffer = (double *)calloc((size_t)plan->m * 2, sizeof(double));
    s->out_buffer = (double *)calloc((size_t)plan->n * 2, sizeof(double));
    s->real_buffer_1 = (double *)calloc((size_t)plan->n, sizeof(double));
    s->real_buffer_2 = (double *)calloc((size_t)plan->n, sizeof(double));
    s->cheb_buffer = (double *)calloc((size_t)kl + 1, sizeof(double));
    s->bessel_buffer_1 = (double *)calloc((size_t)kl + 1, sizeof(double));
    s->bessel_buffer_2 = (double *)calloc((size_t)kl + 1 + (size_t)extra, sizeof(double));
    if (!s->in_buffer || !s->out_buffer || !s->real_buffer_1 || !s->real_buffer_2 ||
        !s->cheb_buffer || !s->bessel_buffer_1 || !s->bessel_buffer_2) {
        nufht_scratch_free(s);
        return -301;
    }
    s->m_alloc = plan->m;
    s->n_alloc = plan->n;
    s->Kloc_alloc = kl;
    s->extra_alloc = extra;
    return 0;
}

void nufht_scratch_free(NufhtScratch *s)
{
    if (!s) return;
    nufht_engine_scratch_release(s);
    free(s->in_buffer);
    free(s->out_buffer);
    free(s->real_buffer_1);
    free(s->real_buffer_2);
    free(s->cheb_buffer);
    free(s->bessel_buffer_1);
    free(s->bessel_buffer_2);
    memset(s, 0, sizeof(*s));
}

int nufht_batch(const NufhtPlan *plan, const double *cs, int ld_cs, double *gs, int ld_gs,
                int batch, NufhtScratch *scratch)
{
    if (!plan || !cs || !gs || batch <= 0) return -400;
    if (ld_cs < plan->m || ld_gs < plan->n) return -401;
    if (!plan->impl) return -400;
    return nufht_engine_batch_host(plan, cs, ld_cs, gs, ld_gs, batch, scratch);
}

void nufht_inplace(double *gs, double nu, const double *rs, int m, const double *cs,
                   const double *ws, int n, const NufhtOptions *opt)
{
    if (!gs || !rs || !cs || !ws || m <= 0 || n <= 0) {
        fprintf(stderr, "nufht_inplace: invalid arguments\n");
        return;
    }
    NufhtPlan plan;
    if (nufht_plan_init(&plan, nu, rs, m, ws, n, opt) != 0) {
        fprintf(stderr, "nufht_inplace: plan init failed\n");
        return;
    }
    memset(gs, 0, sizeof(double) * (size_t)n);
    int rc = nufht_batch(&plan, cs, m, gs, n, 1, NULL);
    if (rc != 0) fprintf(stderr, "nufht_inplace: batch execution failed (%d)\n", rc);
    nufht_plan_free(&plan);
}

double *nufht(double nu, const double *rs, int m, const double *cs, const double *ws, int n,
              const NufhtOptions *opt)
{
    double *gs = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!gs) return NULL;
    nufht_inplace(gs, nu, rs, m, cs, ws, n, opt);
    return gs;
}
