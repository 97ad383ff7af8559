/*
 * Plan tables: Hankel coefficients a_k(nu), the asymptotic/local split
 * z(nu, tol, K) and the local-expansion length K_loc(nu, tol, K).
 *
 * The reference precomputes three global tables once per process and
 * caches them on disk (src/bounds.c:306-324, src/fast_hankel_nufht.c:95-312).
 * Here each entry is computed on demand from the same definitions, and
 * memoised, so planning never pays for the 24k Newton solves it does not
 * need:
 *   a_k(nu)            src/bounds.c:63-82   (DLMF 10.17.1)
 *   asymptotic bound   src/bounds.c:120-131
 *   Newton solve       src/bounds.c:141-178 (x0 = 1, rel. step 1e-8)
 *   Wimp bound / K     src/bounds.c:223-297
 * Deviation: the reference never fills slot K = 10 (its loop stops at
 * K_asy = 10, stored at slot 9, src/bounds.c:202) although planning reads
 * slot K_asy (src/fast_hankel_nufht.c:393), so integer orders whose K_asy
 * clamps to 10 get z_split = 0 and K_loc = 0.  Slot 10 is filled here by the
 * same Newton solve with K = 11.
 *
 * A cache file in the reference's format (src/fast_hankel_nufht.c:186-247),
 * found where the reference looks for it ($NUFHT_TABLE_CACHE, then
 * data/nufht_tables.cache under the working directory or its parent, then
 * next to the executable), supplies its z and K entries instead, which
 * reproduces the reference's boxes bit for bit.  Unlike the reference,
 * planning never writes the file: nufht_amd_write_table_cache does, on
 * request.
 */
#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <sys/stat.h>
#include <unistd.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define T_NU 201
#define T_TOL 12
#define T_K 11

static const double TOLS[T_TOL] = {1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9,
                                   1e-10, 1e-11, 1e-12, 1e-13, 1e-14, 1e-15};

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static double g_z[T_NU][T_TOL][T_K];
static int g_k[T_NU][T_TOL][T_K];
static unsigned char g_have[T_NU][T_TOL][T_K];
static int g_cache_tried = 0;

static double a_coef(int k, double nu)
{
    if (k == 0) return 1.0;
    double prod = 1.0;
    for (int j = 1; j <= 2 * k - 1; j += 2) prod *= 4.0 * nu * nu - (double)(j * j);
    double fact = 1.0;
    for (int j = 2; j <= k; ++j) fact *= (double)j;
    return prod / (fact * pow(8.0, (double)k));
}

double nufht_asy_error_bound(double nu, int K, double z)
{
    double a1 = fabs(a_coef(2 * K, nu)), a2 = fabs(a_coef(2 * K + 1, nu));
    return sqrt(2.0 / (M_PI * z)) *
           (a1 / pow(z, (double)(2 * K)) + a2 / pow(z, (double)(2 * K + 1)));
}

static double solve_split(double nu, int K, double tol)
{
    double x = 1.0;
    for (int it = 0; it < 1000; ++it) {
        double f = nufht_asy_error_bound(nu, K, x) - tol;
        double h = 1e-5 * (fabs(x) > 1.0 ? fabs(x) : 1.0);
        double df = ((nufht_asy_error_bound(nu, K, x + h) - tol) -
                     (nufht_asy_error_bound(nu, K, x - h) - tol)) / (2.0 * h);
        if (fabs(df) < 1e-16) break;
        double dx = -f / df;
        x += dx;
        if (x <= 0.0) x = fabs(x);
        if (fabs(dx / x) <= 1e-8) break;
    }
    return x;
}

static double psi_p(double p)
{
    double r = sqrt(fmax(0.0, 1.0 - p * p));
    return log(p) + r - log(1.0 + r);
}

static double wimp_bound(double nu, int K, double z)
{
    double bK = psi_p(z / (2.0 * K + 2.0 + nu));
    double tkn = 2.0 * K - nu;
    if (tkn > 0.0 && z / tkn <= 1.0) {
        double cK = psi_p(z / (2.0 * K + 2.0 - nu));
        return 2.0 * exp(bK * (nu / 2.0 + K + 1.0) + cK * (-nu / 2.0 + K + 1.0)) /
               (1.0 - exp(bK + cK));
    }
    return 2.0 * exp(bK * (nu / 2.0 + K + 1.0)) / (1.0 - exp(bK));
}

static int wimp_terms(double nu, double z, double tol)
{
    int k = (int)ceil((z - nu) / 2.0);
    if (k < 1) k = 1;
    /* no cap: the K = 1 slots of small tolerances need ~1e5 terms (never
     * used by a plan, whose K_asy is at least 2, but written to the cache) */
    while (!(wimp_bound(nu, k, z) < tol) && k < (1 << 30)) k++;
    return k;
}

/* The cache file to read, in the reference's order of preference
 * (src/fast_hankel_nufht.c:95-160): $NUFHT_TABLE_CACHE; else
 * data/nufht_tables.cache under the working directory, if it has a data/
 * directory, or under its parent, if that has one; else the same name next
 * to the executable.  Returns 0 and the path in out, or -1. */
static int cache_path(char *out, size_t outlen)
{
    const char *env = getenv("NUFHT_TABLE_CACHE");
    if (env) {
        if (!*env || strlen(env) + 1 > outlen) return -1;
        strcpy(out, env);
        return 0;
    }
    const char *fname = "data/nufht_tables.cache";
    char cwd[PATH_MAX];
    struct stat st;
    if (getcwd(cwd, sizeof(cwd))) {
        for (int up = 0; up < 2; ++up) {
            char dir[PATH_MAX];
            size_t len = strlen(cwd);
            if (up) {
                const char *slash = strrchr(cwd, '/');
                if (!slash) break;
                len = (size_t)(slash - cwd);
            }
            if (len + 6 >= sizeof(dir) || len + 1 + strlen(fname) + 1 > outlen) continue;
            memcpy(dir, cwd, len);
            strcpy(dir + len, "/data");
            if (stat(dir, &st) == 0 && S_ISDIR(st.st_mode)) {
                memcpy(out, cwd, len);
                out[len] = '/';
                strcpy(out + len + 1, fname);
                return 0;
            }
        }
    }
    char exe[PATH_MAX];
    ssize_t r = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
    if (r <= 0) return -1;
    exe[r] = '\0';
    const char *slash = strrchr(exe, '/');
    if (!slash) return -1;
    size_t len = (size_t)(slash - exe);
    if (len == 0 || len + 1 + strlen(fname) + 1 > outlen) return -1;
    memcpy(out, exe, len);
    out[len] = '/';
    strcpy(out + len + 1, fname);
    return 0;
}

static void try_cache(void)
{
    char path[PATH_MAX];
    if (cache_path(path, sizeof(path))) return;
    FILE *f = fopen(path, "rb");
    if (!f) return;
    uint32_t hdr[6];
    size_t nz = 401 * 32 * 16;
    double *z = NULL;
    int *k = NULL;
    if (fread(hdr, sizeof(uint32_t), 6, f) != 6 || hdr[0] != 0x4E5546u || hdr[1] != 1u ||
        hdr[2] != 401u || hdr[3] != 401u || hdr[4] != 32u || hdr[5] != 16u)
        goto out;
    if (fseek(f, (long)(sizeof(double) * 401 * 64), SEEK_CUR) != 0) goto out;
    z = (double *)malloc(sizeof(double) * nz);
    k = (int *)malloc(sizeof(int) * nz);
    if (!z || !k) goto out;
    if (fread(z, sizeof(double), nz, f) != nz || fread(k, sizeof(int), nz, f) != nz) goto out;
    for (int i = 0; i < T_NU; ++i)
        for (int j = 0; j < T_TOL; ++j)
            for (int s = 0; s < 10; ++s) {     /* slot 10 is empty in such files */
                size_t at = ((size_t)i * 32 + j) * 16 + s;
                g_z[i][j][s] = z[at];
                g_k[i][j][s] = k[at];
                g_have[i][j][s] = 1;
            }
out:
    free(z);
    free(k);
    fclose(f);
}

int nufht_tables_lookup(int nu_int, int j_idx, int k_slot, double *z, int *wimp_k)
{
    if (nu_int < 0 || nu_int >= T_NU || j_idx < 0 || k_slot < 0 || k_slot >= T_K) return -1;
    double tol = TOLS[j_idx < T_TOL ? j_idx : T_TOL - 1];
    pthread_mutex_lock(&g_lock);
    if (!g_cache_tried) {
        g_cache_tried = 1;
        try_cache();
    }
    int jj = j_idx < T_TOL ? j_idx : T_TOL - 1;
    if (!g_have[nu_int][jj][k_slot]) {
        /* slot s holds the split for K = s + 1 asymptotic terms */
        double zz = solve_split((double)nu_int, k_slot + 1, tol);
        g_z[nu_int][jj][k_slot] = zz;
        g_k[nu_int][jj][k_slot] = wimp_terms((double)nu_int, zz, tol);
        g_have[nu_int][jj][k_slot] = 1;
    }
    *z = g_z[nu_int][jj][k_slot];
    *wimp_k = g_k[nu_int][jj][k_slot];
    pthread_mutex_unlock(&g_lock);
    return 0;
}

int nufht_tables_row(double abs_nu, double *coef64)
{
    for (int J = 0; J < 64; ++J) coef64[J] = (J <= 21) ? a_coef(J, abs_nu) : 0.0;
    return 0;
}

/* Writes every table entry the reference's generator fills (a_k for
 * nu = 0, 1/2, ..., 200 and k <= 21; z and K_loc for integer nu = 0..200,
 * the 12 tolerances and K = 1..10 at slots 0..9) in the reference's cache
 * format (src/fast_hankel_nufht.c:186-247): a header of six uint32
 * {0x4E5546, 1, 401, 401, 32, 16}, then double asy_coef[401][64], double
 * z[401][32][16], int32 K[401][32][16], native endian, every other entry
 * zero (slot 10 too, as the reference leaves it).  Returns 0, -1 on an
 * allocation failure, -2 if the file cannot be written. */
int nufht_amd_write_table_cache(const char *path)
{
    const size_t nz = 401 * 32 * 16;
    double *coef = (double *)calloc(401 * 64, sizeof(double));
    double *z = (double *)calloc(nz, sizeof(double));
    int32_t *k = (int32_t *)calloc(nz, sizeof(int32_t));
    int rc = 0;
    if (!coef || !z || !k) {
        rc = -1;
        goto out;
    }
    for (int i = 0; i < 401; ++i) nufht_tables_row(0.5 * i, coef + (size_t)i * 64);
    for (int nu = 0; nu < T_NU; ++nu)
        for (int j = 0; j < T_TOL; ++j)
            for (int slot = 0; slot < 10; ++slot) {
                double zz;
                int kk;
                nufht_tables_lookup(nu, j, slot, &zz, &kk);
                size_t at = ((size_t)nu * 32 + j) * 16 + slot;
                z[at] = zz;
                k[at] = kk;
            }
    {
        const uint32_t hdr[6] = {0x4E5546u, 1u, 401u, 401u, 32u, 16u};
        FILE *f = path ? fopen(path, "wb") : NULL;
        if (!f) {
            rc = -2;
            goto out;
        }
        if (fwrite(hdr, sizeof(uint32_t), 6, f) != 6 ||
            fwrite(coef, sizeof(double), 401 * 64, f) != 401 * 64 ||
            fwrite(z, sizeof(double), nz, f) != nz || fwrite(k, sizeof(int32_t), nz, f) != nz)
            rc = -2;
        if (fclose(f) != 0) rc = -2;
    }
out:
    free(coef);
    free(z);
    free(k);
    return rc;
}
