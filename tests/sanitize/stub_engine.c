/*
 * Host-only stand-in for the device engine (engine.hip), so that the host C
 * of the library (plan.c, tables.c, pynufht.c) can be built and exercised
 * under AddressSanitizer / UndefinedBehaviorSanitizer on a machine without
 * a GPU.  Execution entry points report a device failure (-510); nothing
 * here is part of the library.
 */
#include <stdarg.h>
#include <stdio.h>

#include "internal.h"

static char g_err[1024];

void nufht_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *nufht_amd_last_error(void) { return g_err; }

void nufht_engine_release(NufhtImplHost *h) { (void)h; }

void nufht_engine_scratch_release(NufhtScratch *s) { (void)s; }

int nufht_engine_batch_host(const NufhtPlan *plan, const double *cs, int ld_cs, double *gs,
                            int ld_gs, int batch, NufhtScratch *scratch)
{
    (void)plan; (void)cs; (void)ld_cs; (void)gs; (void)ld_gs; (void)batch; (void)scratch;
    nufht_set_error("no device in the sanitizer build");
    return -510;
}
