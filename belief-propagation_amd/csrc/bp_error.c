/*
 * bp_error.c -- error reporting for libbp_amd.so.
 *
 * The reference reports every failure through assert() or exit()
 * (src/graph/graph.c:91-92,158-159,184,201; src/cuda_benchmark/
 * belief-propagation.cu:519-527).  The drop-in keeps that behaviour by
 * default: bp_fail() prints the message and aborts.  A caller that prefers
 * status codes calls bp_set_abort_on_error(0); failures then only record the
 * message (per thread) and the entry point returns early.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bp_internal.h"
#include "../../include/bp_hip.h"

static __thread char last_error[512];
static __thread int error_set;
static int abort_on_error = 1;

int bp_set_abort_on_error(int enabled)
{
    int previous = abort_on_error;
    abort_on_error = enabled ? 1 : 0;
    return previous;
}

const char *bp_last_error(void)
{
    return error_set ? last_error : "";
}

void bp_clear_error(void)
{
    error_set = 0;
    last_error[0] = '\0';
}

int bp_has_error(void) { return error_set; }

/* formats through a scratch buffer: callers may pass bp_last_error() itself */
static void record(const char *fmt, va_list ap)
{
    char msg[sizeof(last_error)];
    vsnprintf(msg, sizeof(msg), fmt, ap);
    memcpy(last_error, msg, sizeof(msg));
    error_set = 1;
}

void bp_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    record(fmt, ap);
    va_end(ap);
}

void bp_fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    record(fmt, ap);
    va_end(ap);
    if (abort_on_error) {
        fprintf(stderr, "libbp_amd: %s\n", last_error);
        fflush(stderr);
        abort();
    }
}
