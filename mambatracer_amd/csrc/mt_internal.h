// Internal declarations shared by the libmambart translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/mambart.h"

// ---------------------------------------------------------------- errors
void mt_set_error(const char* fmt, ...);

#define MT_HIP_CHECK(call)                                                   \
    do {                                                                     \
        hipError_t mt_e_ = (call);                                           \
        if (mt_e_ != hipSuccess) {                                           \
            mt_set_error("%s failed: %s (%s:%d)", #call,                     \
                         hipGetErrorString(mt_e_), __FILE__, __LINE__);      \
            return MT_ERR_HIP;                                               \
        }                                                                    \
    } while (0)

#define MT_ARG_CHECK(cond, ...)                                              \
    do {                                                                     \
        if (!(cond)) {                                                       \
            mt_set_error(__VA_ARGS__);                                       \
            return MT_ERR_ARG;                                               \
        }                                                                    \
    } while (0)

#define MT_LAUNCH_CHECK() MT_HIP_CHECK(hipGetLastError())

// Teardown: every release is attempted; the first failure is kept so that a
// destroy call can still report it.
struct MtFirstError {
    hipError_t e = hipSuccess;
    void operator()(hipError_t x)
    {
        if (e == hipSuccess) e = x;
    }
    int32_t status(const char* what) const
    {
        if (e == hipSuccess) return MT_OK;
        mt_set_error("%s: %s", what, hipGetErrorString(e));
        return MT_ERR_HIP;
    }
};

// ---------------------------------------------------------------- objects
struct mt_context {
    int device;
    hipStream_t stream;
    bool own_stream;
    // Device workspace reused by the scan/compaction look-back (tile status
    // words, the tile ticket and the device-side epoch that tags the status
    // words so they never need clearing).  Grown on demand (zero-filled);
    // calls on one context are serialised on its stream, so one workspace
    // suffices.
    void* scratch;
    size_t scratch_bytes;
    // Second workspace for plain per-call temporaries (never the status
    // words above, which must only ever hold epoch-tagged values).
    void* temp;
    size_t temp_bytes;
    // A second stream of the context's own (default priority), created on
    // first use: work forked to it from `stream` through an event is joined
    // back to `stream` through another before the call returns, so callers
    // still see one ordered stream.
    hipStream_t side;
    // the fork and join events of the bounce loop (one pair per bounce),
    // created with the side stream; calls on a context are serialised, so
    // every render reuses them
    hipEvent_t fork[64], join[64];
    int nfj;
};

// The context's side stream and its per-bounce fork / join events (created
// on first use).
int32_t mt_ctx_side_stream(mt_context* ctx, hipStream_t* out);

// Packed traversal layout derived from the reference arrays.
//   node (64 B): float4 a = (L.min.xyz, L.max.x), b = (L.max.yz, R.min.xy),
//                float4 c = (R.min.z, R.max.xyz), int4 d = (left, right, parent, sibling)
//   refs: >= 0 interior node (packed index), < 0 leaf ~rank
//   tri  (48 B): float4 (p1.xyz, e1.x), (e1.yz, e2.xy), (e2.z, prim bits, parent, sibling)
struct mt_scene {
    mt_context* ctx;
    mt_scene_desc desc;
    int32_t num_nodes;     // 2N - 1
    int32_t num_interior;  // N - 1
    float4* pnodes;        // [max(num_interior,1) * 4]
    float4* ptris;         // [num_tris * 3]
    int32_t root_ref;      // 0, or ~0 when the tree is a single leaf
};

struct mt_lightgrid {
    mt_context* ctx;
    float* accum;
    float* func;
    float* sat;
    float cam[3]; // LightGrid.cameraPosition (LightGrid.cl:26), a launch argument
    float* tab;    // per-cell sampling tables [LG_CELLS * 640], rebuilt with the SAT
    bool tab_valid; // false after the raw arrays are handed out, until the next SAT build
    // Accumulation of the render loop's guided environment mode, in fixed
    // point (units of 2^-20 luminance, integer atomics: the per-entry sums do
    // not depend on the order of the additions); folded into accum, and
    // cleared, at the end of every render.  LG_PEND_COPIES copies of
    // LG_SIZE entries.
    unsigned long long* pend;
};

struct mt_envmap {
    mt_context* ctx;
    const float* rgba;
    const float* lum;
    const float* sat;
    int32_t w, h;
    // library-owned sampling tables (see EnvView in mt_sat.h), one allocation
    float* tab;
    float *marg, *cond, *fint, *pmarg;
    int *gmarg, *gcond; // search guides over marg / cond rows (same allocation)
    // per-tile tables of the 16 x 32 light-grid tiles' SAT regions (null when
    // the map is smaller than the tiling or the tables would be too large)
    float* ttab;
    int32_t tsxm, tsym, tstride;
    // light grid of the guided environment mode (mt_env_attach_lightgrid;
    // EnvironmentLight.lightGrid, Light.cl:15-33), or null
    mt_lightgrid* grid;
};

// Guide intervals of the environment sampling tables: the marginal CDF gets
// ENV_GUIDE_M + 2 entries, every conditional row ENV_GUIDE_C + 2.
constexpr int ENV_GUIDE_M = 1024;
constexpr int ENV_GUIDE_C = 256;

int32_t mt_ctx_scratch(mt_context* ctx, size_t bytes, void** out);
int32_t mt_ctx_temp(mt_context* ctx, size_t bytes, void** out);
int32_t mt_scan_reserve(mt_context* ctx, int32_t capacity);
// No-alloc compaction workspace (mt_scan.hip): resolve once per device at
// context creation; look up later.
int32_t mt_noalloc_setup(int device);
int32_t mt_lightgrid_fold(hipStream_t st, mt_lightgrid* g);
void* mt_noalloc_workspace(int device);

// One scene's packed traversal layout, as a row of a per-job table.
struct MtSceneRef {
    const float4* pn;
    const float4* pt;
    int32_t root;
    int32_t pad;
};

// A queue split into nseg segments of seg_cap entries; segment p holds
// counts[p] entries starting at p * seg_cap.  Producers append with one
// atomic per wave on their own segment's counter: a single counter shared by
// every wave of a 33M-slot launch serialises (one atomic per wave on one
// address over 33M threads measured 5.9 ms; over 64 addresses, 0.18 ms).
// Consumers launch over nseg * seg_cap ids and skip the empty tails.
struct MtSegQueue {
    const int32_t* counts; // null: a plain queue of *d_count entries
    int32_t nseg;
    int32_t seg_cap;
};

// Closest hit (prim, t) over a queue that mixes jobs: slot i's scene is
// refs[tag[2*i] >> shift]; entries i < *d_count (clamped to capacity).
// mode MT_BATCH_TRACE_SPLIT: entries below *d_split (the earlier bounces'
// survivors) with wave pools, the rest (new camera rays) one ray per lane.
int32_t mt_trace_jobs(mt_context* ctx, const MtSceneRef* refs, const int32_t* tag, int32_t shift,
                      const float4* o_tmin, const float4* d_tmax, const int32_t* d_count, int32_t capacity,
                      int32_t* prim, float* t, int32_t mode, const int32_t* d_split);
// Any hit over a segmented queue for one scene (traversal mode as
// mt_trace_any; the per-lane refill mode traverses like the short-stack mode
// here); stats as mt_trace_any, indexed by queue entry.
int32_t mt_trace_closest_seg(mt_context* ctx, const mt_scene* scene, const float4* o_tmin, const float4* d_tmax,
                             MtSegQueue seg, int32_t* prim, float* t, int32_t mode);
int32_t mt_trace_any_seg(mt_context* ctx, const mt_scene* scene, const float4* o_tmin, const float4* d_tmax,
                         MtSegQueue seg, int32_t* occluded, int32_t* stats, int32_t mode);

// Light-grid geometry (LightGrid.cl:9-37, CLightGrid.java:47-58)
constexpr int LG_NU = 1600;
constexpr int LG_NV = 1600;
constexpr int LG_RX = 16;
constexpr int LG_RY = 32;
constexpr int LG_CX = LG_NU / LG_RX;  // 100
constexpr int LG_CY = LG_NV / LG_RY;  // 50
constexpr int LG_CELLS = LG_CX * LG_CY;
constexpr int LG_SIZE = LG_NU * LG_NV;  // 2 560 000
// Copies of the light grid's fixed-point accumulation.  Work-group b adds
// into copy b % LG_PEND_COPIES; work-groups are dispatched round-robin over
// the device's 8 XCDs, so each copy takes one XCD's atomics, and additions to
// one entry from different XCDs (every camera ray that misses the scene adds
// to the same few entries) do not queue on one address.  The fold adds the
// copies: integer sums, the same whichever copy an addition went to.
constexpr int LG_PEND_COPIES = 8;

inline int mt_blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return b < 1 ? 1 : (int)b;
}

// Work-groups for a persistent launch: what the device keeps resident at
// once (CUs x occupancy of `kernel`), or fewer when `work` items need fewer.
template <class K>
int mt_persistent_blocks(mt_context* ctx, K kernel, int threads, int64_t work)
{
    static thread_local int cached_dev = -1, cus = 0;
    if (cached_dev != ctx->device) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
            cus = 256;
        cached_dev = ctx->device;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    int64_t need = (work + threads - 1) / threads;
    int64_t cap = (int64_t)cus * per_cu;
    return (int)(need < cap ? (need < 1 ? 1 : need) : cap);
}
