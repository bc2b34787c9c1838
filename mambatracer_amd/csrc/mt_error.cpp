// Error reporting of the C ABI: a thread-local message per failing call
// (mt_last_error), as the reference's JOCL layer raises a CLException with
// the failing call's message.  Host-only (no HIP), so the host-side readers
// and builders link it without the GPU runtime.
#include "../../include/mambart.h"

#include <cstdarg>
#include <cstdio>

static thread_local char g_err[512] = "";

void mt_set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" int32_t mt_last_error(char* buf, int32_t len)
{
    if (!buf || len <= 0) return MT_ERR_ARG;
    snprintf(buf, (size_t)len, "%s", g_err);
    return MT_OK;
}

extern "C" const char* mt_version(void) { return "mambart 0.2 gfx950"; }
