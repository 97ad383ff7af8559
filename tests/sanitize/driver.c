/*
 * Drives the host C of the library (planning, partition, tables, scratch,
 * argument checks) through the C ABI for the sanitizer build: every case
 * the CPU tests reach, plus unsorted inputs, high orders, tiny and empty
 * boxes and the error paths.  Exit status 0 when every return code is the
 * expected one; the sanitizers abort on any memory or undefined-behaviour
 * error.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int fails = 0;
#define EXPECT(cond)                                                         \
    do {                                                                     \
        if (!(cond)) {                                                       \
            fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__,     \
                    __LINE__, #cond);                                        \
            ++fails;                                                         \
        }                                                                    \
    } while (0)

static unsigned long long rng = 0x9E3779B97F4A7C15ull;
static double uni(double lo, double hi)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return lo + (hi - lo) * (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

static int cmp(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

static void plan_case(double nu, double tol, int m, int n, double R, double K, int sorted)
{
    double *r = malloc(sizeof(double) * m), *k = malloc(sizeof(double) * n);
    for (int j = 0; j < m; ++j) r[j] = uni(0.0, R);
    for (int i = 0; i < n; ++i) k[i] = uni(0.0, K);
    if (sorted) {
        qsort(r, m, sizeof(double), cmp);
        qsort(k, n, sizeof(double), cmp);
    }
    NufhtOptions o = nufht_default_options();
    o.tol = tol;
    NufhtPlan p;
    int rc = nufht_plan_init(&p, nu, r, m, k, n, &o);
    EXPECT(rc == 0);
    if (rc == 0) {
        /* the partition tiles the (k, r) index rectangle */
        long long area = 0;
        const BoxVector *v[3] = {&p.boxes.loc, &p.boxes.asy, &p.boxes.dir};
        for (int t = 0; t < 3; ++t)
            for (size_t q = 0; q < v[t]->length; ++q) {
                Box b = v[t]->data[q];
                if (b.i1 >= b.i0 && b.j1 >= b.j0)
                    area += (long long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1);
            }
        EXPECT(area == (long long)m * n);
        NufhtScratch s;
        EXPECT(nufht_scratch_init(&s, &p) == 0);
        double *cs = calloc((size_t)m * 2, sizeof(double)), *gs = calloc((size_t)n * 2, sizeof(double));
        EXPECT(nufht_batch(&p, cs, m, gs, n, 2, &s) == -510);   /* no device */
        EXPECT(nufht_batch(&p, cs, m - 1, gs, n, 2, &s) == -401);
        EXPECT(nufht_batch(&p, cs, m, gs, n - 1, 2, &s) == -401);
        EXPECT(nufht_batch(&p, cs, m, gs, n, 0, &s) == -400);
        EXPECT(nufht_batch(&p, NULL, m, gs, n, 1, &s) == -400);
        nufht_scratch_free(&s);
        free(cs);
        free(gs);
        nufht_plan_free(&p);
    }
    /* generate_boxes / split_box on the same (possibly unsorted) arrays */
    Boxes bx = generate_boxes(r, m, k, n, 12.0, 0, 1000, 14);
    if (bx.dir.length) {
        int is, js;
        split_box(r, m, k, n, &bx.dir.data[0], 12.0, &is, &js);
        EXPECT(is >= bx.dir.data[0].i0 && is <= bx.dir.data[0].i1);
    }
    boxes_free(&bx);
    free(r);
    free(k);
}

int main(int argc, char **argv)
{
    const char *cache = argc > 1 ? argv[1] : NULL;
    plan_case(0.0, 1e-8, 3000, 2000, 100.0, 100.0, 1);
    plan_case(5.0, 1e-8, 5000, 4000, 300.0, 200.0, 0);
    plan_case(1.0, 1e-6, 4000, 4000, 500.0, 100.0, 0);
    plan_case(-3.0, 1e-10, 2000, 2500, 50.0, 400.0, 1);
    plan_case(1.5, 1e-8, 1000, 800, 30.0, 30.0, 0);
    plan_case(9.5, 1e-12, 900, 700, 40.0, 40.0, 1);
    plan_case(40.0, 1e-8, 3000, 3000, 100.0, 100.0, 0);
    plan_case(60.0, 1e-10, 2000, 2000, 100.0, 100.0, 1);
    plan_case(200.0, 1e-15, 500, 500, 300.0, 300.0, 0);
    plan_case(2.0, 1e-8, 1, 7, 10.0, 10.0, 1);
    plan_case(2.0, 1e-8, 9, 1, 10.0, 10.0, 1);
    plan_case(0.0, 1e-4, 50, 100, 10.0, 10.0, 0);
    /* argument and order errors */
    NufhtPlan p;
    double r[3] = {1.0, 2.0, 3.0};
    EXPECT(nufht_plan_init(&p, 0.0, NULL, 0, NULL, 0, NULL) == -200);
    EXPECT(nufht_plan_init(&p, 0.3, r, 3, r, 3, NULL) == -2);
    EXPECT(nufht_plan_init(&p, -1.5, r, 3, r, 3, NULL) == -2);
    EXPECT(nufht_plan_init(&p, 201.0, r, 3, r, 3, NULL) == -2);
    NufhtOptions o = nufht_default_options();
    o.tol = 1e-16;
    EXPECT(nufht_plan_init(&p, 1.0, r, 3, r, 3, &o) == -1);
    /* the legacy entry points print their device error and return */
    double g[3];
    nufht_inplace(g, 0.0, r, 3, r, r, 3, NULL);
    double *gp = nufht(0.0, r, 3, r, r, 3, NULL);
    EXPECT(gp != NULL);
    free(gp);
    /* grid sizes, including the refused ones */
    int L, Nf;
    EXPECT(nufht_asy_grid(500.0, M_PI / 1000.0, 10, &L, &Nf) == 0);
    EXPECT(nufht_asy_grid(1e6, M_PI / 2e6, 10, &L, &Nf) == -541);
    /* tables: every order and tolerance index a plan can ask for */
    for (int nu = 0; nu <= 200; nu += 13)
        for (int j = 0; j < 12; j += 3)
            for (int s = 0; s <= 10; s += 5) {
                double z;
                int kk;
                EXPECT(nufht_tables_lookup(nu, j, s, &z, &kk) == 0);
            }
    if (cache) EXPECT(nufht_amd_write_table_cache(cache) == 0);
    EXPECT(nufht_amd_write_table_cache("/nonexistent-dir/x.cache") == -2);
    if (fails) fprintf(stderr, "%d expectation(s) failed\n", fails);
    else printf("sanitizer driver: all cases ran clean\n");
    return fails ? 1 : 0;
}
