/*
 * oracle/plan.c -- TEST INFRASTRUCTURE (the oracle), never product code.
 *
 * Literal restatement of the reference planner, keeping its linear
 * "first index greater than" scans:
 *   fill_plan_parameters   src/fast_hankel_nufht.c:322-414
 *   find_first_*           src/fast_hankel_nufht.c:454-466
 *   splitting_objective    src/fast_hankel_nufht.c:498-521
 *   split_box              src/fast_hankel_nufht.c:523-574
 *   generate_boxes         src/fast_hankel_nufht.c:578-744
 * Pinned against oracle/_ref (the reference planner compiled from its own
 * sources) through tests/golden/ref_boxes.json.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double *g_coef, *g_z;
static int *g_k;
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void tables_init(void)
{
    g_coef = (double *)malloc(sizeof(double) * ORC_NU_ROWS * ORC_ASY_COEF);
    g_z = (double *)malloc(sizeof(double) * ORC_TI * ORC_TJ * ORC_TK);
    g_k = (int *)malloc(sizeof(int) * ORC_TI * ORC_TJ * ORC_TK);
    orc_generate_tables(g_coef, g_z, g_k);
}

static int is_int(double x) { return fabs(x - nearbyint(x)) < 1e-12; }

int orc_fill_params(OrcParams *p, double nu, double tol, int max_levels,
                    int min_dim_prod, double z_split, int K_asy, int K_loc)
{
    memset(p, 0, sizeof(*p));
    p->max_levels = max_levels;
    p->min_dim_prod = min_dim_prod;
    if (tol < 1e-15) return -1;
    int two_int = is_int(2.0 * nu), nu_int = is_int(nu);
    if (!two_int || (nu_int && fabs(nu) > 200.0) ||
        (!nu_int && two_int && fabs(nu) > 9.5))
        return -2;
    pthread_once(&g_once, tables_init);
    int i_idx = (int)llround(fabs(nu));
    if (i_idx < 0 || i_idx >= ORC_TI) return -3;
    int row = (int)llround(2.0 * fabs(nu) + 1.0) - 1;
    if (row < 0 || row >= ORC_NU_ROWS) return -4;
    memcpy(p->asy_coef, g_coef + (size_t)row * ORC_ASY_COEF, sizeof(double) * 64);
    int j = (int)ceil(-log10(tol) - 3.0);
    if (j < 1) j = 1;
    int j_idx = j - 1;
    if (j_idx < 0 || j_idx >= ORC_TJ) return -5;
    p->nu = nu;
    p->tol = tol;
    if (nu_int) {
        int k;
        if (K_asy <= 0) {
            int k0 = (int)floor(fabs(nu) / 5.0 + log10(1.0 / tol) / 4.0 + 1.0);
            k = k0 < 10 ? k0 : 10;
        } else {
            k = K_asy;
        }
        if (k < 0 || k >= ORC_TK) return -6;
        p->K_asy = k;
        size_t at = ((size_t)i_idx * ORC_TJ + j_idx) * ORC_TK + k;
        p->z_split = isnan(z_split) ? g_z[at] : z_split;
        p->K_loc = K_loc <= 0 ? g_k[at] : K_loc;
    } else {
        p->K_asy = (int)llround(fabs(nu) - 0.5);
        p->z_split = NAN;
        p->K_loc = -1;
    }
    return 0;
}

static int first_greater(const double *v, int n, double t)
{
    for (int i = 0; i < n; ++i)
        if (v[i] > t) return i;
    return -1;
}

typedef struct { OrcBox *d; int n, cap; } Vec;

static void push(Vec *v, int i0, int i1, int j0, int j1)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 8;
        v->d = (OrcBox *)realloc(v->d, sizeof(OrcBox) * (size_t)v->cap);
    }
    OrcBox b = {i0, i1, j0, j1};
    v->d[v->n++] = b;
}

static long long objective(const double *rs, int m, const double *ws,
                           const OrcBox *b, double z, int i)
{
    int j = first_greater(rs, m, z / ws[i]);
    if (j < 0) return 0;
    return (long long)(i - b->i0 + 1) * (long long)(j - b->j0 + 1) +
           (long long)(b->i1 - i + 1) * (long long)(b->j1 - j + 1);
}

void orc_split_box(const double *rs, int m, const double *ws, int n,
                   const OrcBox *box, double z, int *ispl, int *jspl)
{
    (void)n;
    int cand[10], nc = 0;
    if (box->i1 <= box->i0) {
        cand[nc++] = box->i0;
    } else {
        for (int k = 0; k < 10; ++k) {
            double t = (double)k / 9.0;
            int idx = (int)llround(box->i0 + t * (double)(box->i1 - box->i0));
            if (idx < box->i0) idx = box->i0;
            if (idx > box->i1) idx = box->i1;
            if (nc == 0 || idx != cand[nc - 1]) cand[nc++] = idx;
        }
    }
    long long best = -1;
    int bi = box->i0;
    for (int c = 0; c < nc; ++c) {
        long long v = objective(rs, m, ws, box, z, cand[c]);
        if (v > best) { best = v; bi = cand[c]; }
    }
    int j = first_greater(rs, m, z / ws[bi]);
    if (j < 0) j = box->j0;
    *ispl = bi;
    *jspl = j;
}

int orc_generate_boxes(const double *rs, int m, const double *ws, int n,
                       double z, int max_levels, int min_dim_prod, int kloc,
                       OrcBoxes *out)
{
    Vec loc = {0}, asy = {0}, dir = {0};
    memset(out, 0, sizeof(*out));
    if (n <= 0 || m <= 0) return 0;
    if (max_levels <= 0) {
        int mn = n < m ? n : m;
        double ratio = (double)mn * (double)mn / (double)min_dim_prod;
        max_levels = ratio < 1.0 ? 0 : (int)floor(log2(ratio));
        if (max_levels < 0) max_levels = 0;
    }
    int i0 = first_greater(ws, n, z / rs[m - 1]);
    int i1, j0, j1;
    if (m == 1) {
        i1 = i0 == -1 ? -1 : i0 - 1;
    } else {
        int idx = first_greater(ws, n, z / rs[0]);
        i1 = idx == -1 ? n - 1 : idx - 1;
    }
    j0 = first_greater(rs, m, z / ws[n - 1]);
    if (n == 1) {
        j1 = j0 == -1 ? -1 : j0 - 1;
    } else {
        int idx = first_greater(rs, m, z / ws[0]);
        j1 = idx == -1 ? m - 1 : idx - 1;
    }
    if (i0 == -1) {
        push(&loc, 0, n - 1, 0, m - 1);
        goto done;
    } else if (i1 < 0) {
        push(&asy, 0, n - 1, 0, m - 1);
        goto done;
    }
    if (i0 > 0) push(&loc, 0, i0 - 1, 0, m - 1);
    if (j0 > 0) push(&loc, i0, n - 1, 0, j0 - 1);
    if (i1 < n - 1) push(&asy, i1 + 1, n - 1, 0, m - 1);
    if (j1 < m - 1) push(&asy, 0, i1, j1 + 1, m - 1);
    if (m == 1 || n == 1) goto done;
    push(&dir, i0, i1, j0, j1);
    for (int level = 0; level < max_levels; ++level) {
        Vec nd = {0};
        for (int bi = 0; bi < dir.n; ++bi) {
            OrcBox b = dir.d[bi];
            long long d1 = b.i1 - b.i0 + 1, d2 = b.j1 - b.j0 + 1;
            if (d1 * d2 > 4LL * (long long)min_dim_prod) {
                int is, js;
                orc_split_box(rs, m, ws, n, &b, z, &is, &js);
                if (is > b.i0) {
                    push(&loc, b.i0, is - 1, b.j0, js - 1);
                    push(&nd, b.i0, is - 1, js, b.j1);
                }
                push(&asy, is, b.i1, js, b.j1);
                push(&nd, is, b.i1, b.j0, js - 1);
            } else {
                push(&nd, b.i0, b.i1, b.j0, b.j1);
            }
        }
        free(dir.d);
        dir = nd;
    }
    {
        Vec nl = {0}, na = {0};
        for (int bi = 0; bi < loc.n; ++bi) {
            OrcBox b = loc.d[bi];
            int d1 = b.i1 - b.i0 + 1, d2 = b.j1 - b.j0 + 1;
            if ((d1 < d2 ? d1 : d2) < kloc) push(&dir, b.i0, b.i1, b.j0, b.j1);
            else push(&nl, b.i0, b.i1, b.j0, b.j1);
        }
        for (int bi = 0; bi < asy.n; ++bi) {
            OrcBox b = asy.d[bi];
            long long area = (long long)(b.i1 - b.i0 + 1) * (long long)(b.j1 - b.j0 + 1);
            if (area < (long long)min_dim_prod) push(&dir, b.i0, b.i1, b.j0, b.j1);
            else push(&na, b.i0, b.i1, b.j0, b.j1);
        }
        free(loc.d);
        free(asy.d);
        loc = nl;
        asy = na;
    }
done:
    out->loc = loc.d; out->nloc = loc.n;
    out->asy = asy.d; out->nasy = asy.n;
    out->dir = dir.d; out->ndir = dir.n;
    return 0;
}

void orc_boxes_free(OrcBoxes *b)
{
    free(b->loc);
    free(b->asy);
    free(b->dir);
    memset(b, 0, sizeof(*b));
}
