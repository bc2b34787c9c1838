// Wavefront bounce loop of one job: restates CDeviceGI.loop
// (src/cl/device/CDeviceGI.java:214-271); the per-path work of a bounce is
// shade_path (mt_path.h).
//
// Layout.  Path state is SoA indexed by path id (pid = sample * pixels +
// pixel): throughput.xyz + last BSDF pdf in one float4, accumulated
// radiance in another.  The live paths of a bounce form a queue of pids with
// their rays (two float4 streams) and a device-resident count; every stage
// reads the count on the device (the reference's `global int* count`), so
// the host enqueues the whole render without synchronising.
//
// Per bounce:
//   1. closest hit over the queue (mt_trace_closest);
//   2. one fused shading kernel per slot (shade_path): hit shading,
//      setupBsdf, LightHitPass, DirectLight's light sample + EvaluateBsdf +
//      shadow-ray setup, SampleBsdf for the next ray and Russian roulette;
//      shadow rays are appended to a segmented queue (one atomic per wave on
//      the counter of the work-group's segment), so the any-hit launch has
//      empty lanes only at segment tails and no counter is contended;
//   3. any hit over the shadow queue (mt_trace_any_seg);
//   4. unoccluded direct light added to the path radiance;
//   5. stable compaction of live slots (pid and next ray) into the other
//      queue buffer (mt_compact; CCompact.execute in the reference).
// After the last bounce the per-path radiance is averaged per pixel in a
// fixed order, so the image is deterministic: no float atomics.
#include "mt_path.h"

#include <vector>

namespace {

constexpr int RT = 256; // threads per block

struct PathState {
    float4* T;   // throughput.xyz, last BSDF pdf (lastPdfW)
    float4* L;   // accumulated radiance
};

__global__ __launch_bounds__(RT) void k_camera(Camera C, int npix, int npaths, uint32_t seed,
                                               float4* __restrict__ o_tmin, float4* __restrict__ d_tmax,
                                               int* __restrict__ queue, PathState S, int* __restrict__ count0)
{
    int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid == 0) *count0 = npaths;
    if (pid >= npaths) return;
    float4 o, d;
    camera_ray(C, pid, npix, seed, o, d);
    o_tmin[pid] = o;
    d_tmax[pid] = d;
    queue[pid] = pid;
    S.T[pid] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);         // InitPathData, Initialize.cl:25-48
    S.L[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

struct ShadeArgs {
    SceneView V;
    PathState S;
    const int* queue;
    const int* d_count;
    float4* o_tmin;       // in: this bounce's rays; out: next rays (same slot)
    float4* d_tmax;
    const int* prim;
    const float* tin;
    int* alive;           // out: 1 if the slot's path continues
    float4* s_o_tmin;     // out: segmented queue of shadow rays
    float4* s_d_tmax;
    float4* s_contrib;    // out: unoccluded contribution, w = the ray's slot (int bits)
    int* shadow_count;    // SHADOW_SEGS segment lengths
    int seg_cap;
    int bounce;
    int rr_start;
    int flags;
    uint32_t seed;
};

// One slot of the fused shading stage; its shadow ray, if any, goes to the
// next free entry of its work-group's shadow-queue segment.
__device__ __forceinline__ void shade_slot(const ShadeArgs& A, int slot, bool live)
{
    PathOut P;
    P.has_shadow = false;
    if (live) {
        const int pid = A.queue[slot];
        shade_path(A.V, A.seed, A.rr_start, A.flags, pid, A.bounce, A.o_tmin[slot], A.d_tmax[slot], A.S.T[pid],
                   A.prim[slot], A.tin[slot], P);
        if (P.alive) {
            A.S.T[pid] = P.T;
            A.o_tmin[slot] = P.ro;
            A.d_tmax[slot] = P.rd;
        }
        if (P.Ladd.x != 0.0f || P.Ladd.y != 0.0f || P.Ladd.z != 0.0f) {
            float4 L = A.S.L[pid];
            A.S.L[pid] = make_float4(L.x + P.Ladd.x, L.y + P.Ladd.y, L.z + P.Ladd.z, L.w);
        }
        A.alive[slot] = P.alive ? 1 : 0;
    }
    const int seg = blockIdx.x % SHADOW_SEGS;
    const int k = seg * A.seg_cap + wave_append(P.has_shadow, A.shadow_count + seg);
    if (P.has_shadow) {
        A.s_o_tmin[k] = P.so;
        A.s_d_tmax[k] = P.sd;
        A.s_contrib[k] = make_float4(P.contrib.x, P.contrib.y, P.contrib.z, __int_as_float(slot));
    }
}

// One slot per lane over a capacity-sized grid (the segment capacities
// assume exactly that); work-groups past *d_count leave at once.
__global__ __launch_bounds__(RT) void k_shade(ShadeArgs A)
{
    const int count = *A.d_count;
    if ((int64_t)blockIdx.x * blockDim.x >= count) return;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    shade_slot(A, slot, slot < count);
}

// Unoccluded shadow-queue entries add their contribution to their path.
__global__ __launch_bounds__(RT) void k_nee_accum(const int* __restrict__ queue, const int* __restrict__ seg_counts,
                                                  int seg_cap, const float4* __restrict__ contrib,
                                                  const int* __restrict__ occluded, PathState S)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (!seg_live(seg_counts, seg_cap, k) || occluded[k]) return;
    const float4 c = contrib[k];
    const int pid = queue[__float_as_int(c.w)];
    float4 L = S.L[pid];
    S.L[pid] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
}

__global__ void k_zero_i32(int* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Sums per-entry traversal counters {I, L, P} of the live entries: i <
// *d_count, or (seg_counts set) the live entries of a segmented queue.
__global__ __launch_bounds__(RT) void k_sum_ilp(const int* __restrict__ st, const int* __restrict__ d_count,
                                                const int* __restrict__ seg_counts, int seg_cap,
                                                unsigned long long* __restrict__ tot)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long a = 0, b = 0, c = 0;
    if (seg_counts ? seg_live(seg_counts, seg_cap, i) : i < *d_count) {
        a = (unsigned)st[3 * i];
        b = (unsigned)st[3 * i + 1];
        c = (unsigned)st[3 * i + 2];
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        c += __shfl_xor(c, o);
    }
    if (__lane_id() == 0 && (a | b | c)) {
        atomicAdd(&tot[0], a);
        atomicAdd(&tot[1], b);
        atomicAdd(&tot[2], c);
    }
}

} // namespace

extern "C" int32_t mt_render(mt_context* ctx, const mt_scene* scene, const float* materials,
                             const mt_envmap* env, const int32_t* light_faces, const int32_t* d_num_lights,
                             const mt_render_desc* desc, float* image, mt_render_stats* stats)
{
    MT_ARG_CHECK(ctx && scene && materials && desc && image, "mt_render: NULL argument");
    MT_ARG_CHECK(desc->width > 0 && desc->height > 0 && desc->spp > 0, "mt_render: empty image");
    MT_ARG_CHECK(desc->max_bounces >= 1 && desc->max_bounces <= 63, "mt_render: max_bounces %d not in [1,63]",
                 desc->max_bounces);
    MT_ARG_CHECK(!light_faces == !d_num_lights, "mt_render: light_faces and d_num_lights go together");
    int64_t npix = (int64_t)desc->width * desc->height;
    int64_t np64 = npix * desc->spp;
    MT_ARG_CHECK(np64 <= (1ll << 30), "mt_render: %lld paths exceed the 2^30 wavefront limit", (long long)np64);
    int np = (int)np64;
    hipStream_t st = ctx->stream;

    Camera C;
    int32_t rc = camera_from_desc(desc, C);
    if (rc) return rc;

    // ---- buffers (one allocation)
    const int nb = desc->max_bounces;
    size_t n = (size_t)np;
    size_t bytes = 0;
    auto take = [&](size_t b) { size_t off = bytes; bytes += (b + 255) & ~(size_t)255; return off; };
    size_t oT = take(16 * n), oL = take(16 * n);
    size_t oA_o = take(16 * n), oA_d = take(16 * n), oB_o = take(16 * n), oB_d = take(16 * n);
    size_t oQA = take(4 * n), oQB = take(4 * n);
    // shadow rays: a segmented queue filled by k_shade's work-groups
    const int seg_cap = shadow_seg_cap(np, RT);
    const size_t sq = (size_t)SHADOW_SEGS * seg_cap;
    size_t oPrim = take(4 * n), oT_ = take(4 * n), oAlive = take(4 * n), oOcc = take(4 * sq);
    size_t oSo = take(16 * sq), oSd = take(16 * sq), oSc = take(16 * sq);
    size_t oCnt = take(4 * 64), oShc = take(4 * 64 * (size_t)SHADOW_SEGS);
    const bool want_ilp = (desc->flags & MT_RENDER_TRAVERSAL_STATS) != 0;
    const bool profile = (desc->flags & MT_RENDER_PROFILE) != 0;
    size_t oIlp = want_ilp ? take(12 * (sq > n ? sq : n)) : 0, oTot = take(8 * 8);
    char* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) {
        mt_set_error("mt_render: cannot allocate %.2f GB of wavefront state", bytes / 1e9);
        return MT_ERR_ALLOC;
    }
    auto F4 = [&](size_t off) { return reinterpret_cast<float4*>(base + off); };
    auto I = [&](size_t off) { return reinterpret_cast<int*>(base + off); };
    PathState S{F4(oT), F4(oL)};
    float4* ro[2] = {F4(oA_o), F4(oB_o)};
    float4* rd[2] = {F4(oA_d), F4(oB_d)};
    int* q[2] = {I(oQA), I(oQB)};
    int* counts = I(oCnt);
    int* shc = I(oShc);
    auto chk = [&](void) -> int32_t {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            mt_set_error("mt_render: launch failed: %s", hipGetErrorString(e));
            return MT_ERR_HIP;
        }
        return MT_OK;
    };

    SceneView V;
    V.points = reinterpret_cast<const float4*>(scene->desc.points);
    V.normals = reinterpret_cast<const float4*>(scene->desc.normals);
    V.faces = scene->desc.faces;
    V.mats = reinterpret_cast<const MtMaterial*>(materials);
    V.light_faces = light_faces;
    V.num_lights = d_num_lights;
    V.env = env_view(env);

    int grid = mt_blocks(np, RT);
    const int gshade = grid;              // one slot per lane; empty work-groups retire at once
    const int gshadow = mt_blocks((int64_t)sq, RT);
    unsigned long long* tot = reinterpret_cast<unsigned long long*>(base + oTot);
    int* ilp = want_ilp ? I(oIlp) : nullptr;
    hipEvent_t ev[4 * 64];
    int nev = 0;
    if (profile) {
        for (; nev < 4 * nb; nev++)
            if (hipEventCreate(&ev[nev]) != hipSuccess) {
                rc = MT_ERR_HIP;
                mt_set_error("mt_render: hipEventCreate failed");
                break;
            }
    }
    k_zero_i32<<<1, 64, 0, st>>>(counts, 64);
    k_zero_i32<<<mt_blocks(64 * SHADOW_SEGS, RT), RT, 0, st>>>(shc, 64 * SHADOW_SEGS);
    k_zero_i32<<<1, 64, 0, st>>>(reinterpret_cast<int*>(tot), 16);
    k_camera<<<grid, RT, 0, st>>>(C, (int)npix, np, desc->seed, ro[0], rd[0], q[0], S, counts);
    if (rc == MT_OK) rc = chk();
    int cur = 0;
    for (int b = 0; b < nb && rc == MT_OK; b++) {
        int* dcount = counts + b;
        if (profile) hipEventRecord(ev[4 * b], st);
        rc = mt_trace_closest(ctx, scene, reinterpret_cast<float*>(ro[cur]), reinterpret_cast<float*>(rd[cur]),
                              dcount, np, I(oPrim), reinterpret_cast<float*>(base + oT_), ilp, desc->mode);
        if (rc) break;
        if (profile) hipEventRecord(ev[4 * b + 1], st);
        if (want_ilp) k_sum_ilp<<<grid, RT, 0, st>>>(ilp, dcount, nullptr, 0, tot);
        ShadeArgs A;
        A.V = V;
        A.S = S;
        A.queue = q[cur];
        A.d_count = dcount;
        A.o_tmin = ro[cur];
        A.d_tmax = rd[cur];
        A.prim = I(oPrim);
        A.tin = reinterpret_cast<const float*>(base + oT_);
        A.alive = I(oAlive);
        A.s_o_tmin = F4(oSo);
        A.s_d_tmax = F4(oSd);
        A.s_contrib = F4(oSc);
        int* seg_counts = shc + (size_t)b * SHADOW_SEGS;
        A.shadow_count = seg_counts;
        A.seg_cap = seg_cap;
        A.bounce = b;
        A.rr_start = desc->rr_start;
        A.flags = desc->flags;
        A.seed = desc->seed;
        k_shade<<<gshade, RT, 0, st>>>(A);
        if ((rc = chk())) break;
        if (profile) hipEventRecord(ev[4 * b + 2], st);
        // any hit over the segmented shadow queue (segment lengths on the device)
        rc = mt_trace_any_seg(ctx, scene, F4(oSo), F4(oSd), MtSegQueue{seg_counts, SHADOW_SEGS, seg_cap}, I(oOcc),
                              ilp, desc->mode);
        if (rc) break;
        if (profile) hipEventRecord(ev[4 * b + 3], st);
        if (want_ilp) k_sum_ilp<<<gshadow, RT, 0, st>>>(ilp, nullptr, seg_counts, seg_cap, tot + 3);
        k_nee_accum<<<gshadow, RT, 0, st>>>(q[cur], seg_counts, seg_cap, F4(oSc), I(oOcc), S);
        if ((rc = chk())) break;
        if (b + 1 < nb) {
            mt_payload P[3] = {{q[cur], q[1 - cur], 4}, {ro[cur], ro[1 - cur], 16}, {rd[cur], rd[1 - cur], 16}};
            rc = mt_compact(ctx, I(oAlive), 0, dcount, np, nullptr, counts + b + 1, P, 3);
            if (rc) break;
            cur = 1 - cur;
        }
    }
    if (rc == MT_OK) {
        k_resolve<<<mt_blocks(npix, RT), RT, 0, st>>>(S.L, (int)npix, desc->spp, reinterpret_cast<float4*>(image));
        rc = chk();
    }
    if (rc == MT_OK && stats) {
        int hc[64];
        std::vector<int> hs((size_t)nb * SHADOW_SEGS);
        unsigned long long ht[6];
        hipError_t e = hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hs.data(), shc, hs.size() * sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(ht, tot, sizeof(ht), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            mt_set_error("mt_render: stats readback failed: %s", hipGetErrorString(e));
            rc = MT_ERR_HIP;
        } else {
            stats->extension_rays = 0;
            stats->shadow_rays = 0;
            stats->paths = np;
            for (int b = 0; b < 64; b++) stats->bounce_counts[b] = b < nb ? hc[b] : 0;
            for (int b = 0; b < nb; b++) stats->extension_rays += hc[b];
            for (int v : hs) stats->shadow_rays += v;
            for (int k = 0; k < 3; k++) {
                stats->closest_ilp[k] = (int64_t)ht[k];
                stats->any_ilp[k] = (int64_t)ht[3 + k];
            }
            stats->closest_launches = stats->any_launches = 0;
            stats->closest_ms = stats->any_ms = 0.0f;
            if (profile) {
                for (int b = 0; b < nb; b++) {
                    float ms = 0.0f;
                    if (hipEventElapsedTime(&ms, ev[4 * b], ev[4 * b + 1]) == hipSuccess) {
                        stats->closest_ms += ms;
                        stats->closest_launches++;
                    }
                    if (hipEventElapsedTime(&ms, ev[4 * b + 2], ev[4 * b + 3]) == hipSuccess) {
                        stats->any_ms += ms;
                        stats->any_launches++;
                    }
                }
            }
        }
    }
    hipStreamSynchronize(st);
    for (int k = 0; k < nev; k++) hipEventDestroy(ev[k]);
    hipFree(base);
    if (rc == MT_OK) rc = mt_scan_check(ctx);
    return rc;
}
