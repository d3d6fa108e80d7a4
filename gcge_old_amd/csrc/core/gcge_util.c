/*
 * Utilities of the GCG core: timer and printing.
 *
 * GCGE_GetTime follows the role of the reference timer
 * (src/gcge_utilities.c:26-39) but returns wall-clock seconds from
 * CLOCK_MONOTONIC: the reference reads getrusage user CPU time, which does
 * not see time spent waiting for a GPU.
 */
#include <stdarg.h>
#include <stdio.h>
#include <time.h>

#include "gcge.h"

GCGE_DOUBLE GCGE_GetTime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (GCGE_DOUBLE)ts.tv_sec + 1e-9 * (GCGE_DOUBLE)ts.tv_nsec;
}

/* Single-process build: every message is printed (reference
 * src/gcge_utilities.c:95-112 prints on rank 0 only). */
void GCGE_Printf(const char *fmt, ...)
{
    va_list vp;
    va_start(vp, fmt);
    vprintf(fmt, vp);
    va_end(vp);
}
