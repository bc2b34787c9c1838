/*
 * ORACLE (test infrastructure only): evaluates the shared transcendental
 * functions (mt_xmath.h, compiled here for the host) over arrays, so tests
 * can pin them against float64 references.
 *   kind 0 sin(a), 1 cos(a), 2 acos(a), 3 atan2(a, b), 4 a^5
 */
#include "orc.h"
#include "orc_math.h"

int orc_xm_eval(int32_t kind, int64_t n, const float* a, const float* b, float* out)
{
    if (kind < 0 || kind > 4 || n < 0) return ORC_ERR_ARG;
    for (int64_t i = 0; i < n; i++) {
        switch (kind) {
        case 0: out[i] = mt_xm_sinf(a[i]); break;
        case 1: out[i] = mt_xm_cosf(a[i]); break;
        case 2: out[i] = mt_xm_acosf(a[i]); break;
        case 3: out[i] = mt_xm_atan2f(a[i], b[i]); break;
        default: out[i] = mt_xm_pow5f(a[i]); break;
        }
    }
    return ORC_OK;
}
