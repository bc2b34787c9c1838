/*
 * ORACLE (test infrastructure only) — stage D: exclusive scan and stable
 * stream compaction.
 *
 * orc_scan_excl_i32 restates the three-phase PrefixSumInteger pipeline
 * (src/cl/algorithms/PrefixSumInteger.java:80-98; kernels Util.cl:917-980):
 * out[i] = sum(in[0..i-1]), total = in[n-1] + out[n-1] (globalTotalInteger,
 * Util.cl:977-980).  Integer sums wrap modulo 2^32 as the OpenCL int does.
 * Deviation: localScanInteger converts inclusive to exclusive with
 * `aux[l] > 0 ? aux[l-1] : aux[0]` (Util.cl:941-944), which reads aux[-1] for
 * lane 0 of every work-group whose first input is positive and is only an
 * exclusive scan for non-negative inputs; the oracle returns the exclusive
 * scan that PrefixSumIntegerTest expects (all-ones input -> out[i] = i,
 * total = n, test/algorithm/PrefixSumIntegerTest.java:30-36).
 * Pinned by: that all-ones known answer, brute force numpy cumsum on random
 * inputs, and the total == popcount invariant.
 *
 * orc_compact_stable restates identifyMeshLights/prepareLightInfo
 * (LightConfiguration.cl:1-55): out[scan[i]] = i for every i with pred[i],
 * count = total.  CompactAtomic (Compact.cl:21-40) produces the same multiset
 * in an unspecified order; the stable order is the one the build pins.
 * Pinned by: numpy flatnonzero on the same predicate.
 */
#include "orc.h"

int orc_scan_excl_i32(const int32_t* in, int64_t n, int32_t* out, int32_t* total)
{
    if (n < 0 || (n > 0 && (!in || !out))) return ORC_ERR_ARG;
    uint32_t acc = 0;
    for (int64_t i = 0; i < n; i++) {
        out[i] = (int32_t)acc;
        acc += (uint32_t)in[i];
    }
    if (total) *total = n > 0 ? (int32_t)((uint32_t)in[n - 1] + (uint32_t)out[n - 1]) : 0;
    return ORC_OK;
}

int orc_compact_stable(const int32_t* pred, int64_t n, int32_t* out_idx, int32_t* count)
{
    if (n < 0 || (n > 0 && (!pred || !out_idx))) return ORC_ERR_ARG;
    int32_t k = 0;
    for (int64_t i = 0; i < n; i++)
        if (pred[i]) out_idx[k++] = (int32_t)i;
    if (count) *count = k;
    return ORC_OK;
}
