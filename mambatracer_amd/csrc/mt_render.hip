// Wavefront bounce loop of one job: restates CDeviceGI.loop
// (src/cl/device/CDeviceGI.java:214-271); the per-path work of a bounce is
// shade_path (mt_path.h).
//
// Layout.  Path state is SoA indexed by path id (pid = pixel rank * spp +
// sample, pixels ranked in 8x8 tiles: path_pixel in mt_path.h; bounce 0 puts
// path pid in queue slot pid): throughput.xyz + last BSDF pdf in one float4,
// accumulated radiance in another.  The live paths of a bounce form a queue
// of pids with their rays (two float4 streams) and a device-resident count;
// every stage reads the count on the device (the reference's `global int*
// count`), so the host enqueues the whole render without synchronising.
//
// Per bounce:
//   1. closest hit over the queue (mt_trace_closest; at bounce 0
//      k_trace_camera generates each camera ray in place instead);
//   2. one fused shading kernel per slot (shade_path): hit shading,
//      setupBsdf, LightHitPass, DirectLight's light sample + EvaluateBsdf +
//      shadow-ray setup, SampleBsdf for the next ray and Russian roulette;
//      shadow rays are appended to a segmented queue (one atomic per wave on
//      the counter of the work-group's segment), so the any-hit launch has
//      empty lanes only at segment tails and no counter is contended;
//   3. any hit over the shadow queue, an unoccluded ray adding its light to
//      its path's radiance in the same kernel (k_trace_shadow);
//   4. stable compaction of live slots (pid and next ray) into the other
//      queue buffer (mt_compact; CCompact.execute in the reference).
// With traversal counters or a non-default traversal layout the stages run
// unfused (k_camera, mt_trace_any_seg, k_nee_accum): same paths, same bits.
// After the last bounce the per-path radiance is averaged per pixel in a
// fixed order, so the image is deterministic: no float atomics.
#include "mt_path.h"
#include "mt_traverse.h"

#include <vector>

namespace {

constexpr int RT = 256; // threads per block

struct PathState {
    float4* T;   // throughput.xyz, last BSDF pdf (lastPdfW)
    float4* L;   // accumulated radiance
};

// The frame seed is read from device memory so that a captured render can
// be replayed with a new seed (set by k_set_seed before the replay).
__global__ void k_set_seed(uint32_t* d_seed, uint32_t seed)
{
    if (threadIdx.x == 0) *d_seed = seed;
}

__global__ __launch_bounds__(RT) void k_camera(Camera C, int npaths, const uint32_t* __restrict__ d_seed,
                                               float4* __restrict__ o_tmin, float4* __restrict__ d_tmax,
                                               int* __restrict__ queue, PathState S, int* __restrict__ count0)
{
    int pid = blockIdx.x * blockDim.x + threadIdx.x; // slot = path id (see path_pixel)
    if (pid == 0) *count0 = npaths;
    if (pid >= npaths) return;
    const uint32_t seed = *d_seed;
    float4 o, d;
    camera_ray(C, pid, seed, o, d);
    o_tmin[pid] = o;
    d_tmax[pid] = d;
    queue[pid] = pid;
    S.T[pid] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);         // InitPathData, Initialize.cl:25-48
    S.L[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Bounce 0 of the fused loop: path pid = slot traces its camera ray
// (generated here, never stored) to its closest hit; same per-ray traversal
// as mt_trace_closest's default mode, so prim/t are the same bits.  POOL 0:
// one path per lane, work-groups renumbered in per-XCD runs
// (xcd_chunked_group); POOL > 0: each wave traces the run of 64 * POOL
// consecutive paths (4 runs per work-group), refilling finished lanes
// (pool_trace).
template <int POOL>
__global__ __launch_bounds__(RT) void k_trace_camera(Camera C, int npaths, const uint32_t* __restrict__ d_seed,
                                                     const float4* __restrict__ pn, const float4* __restrict__ pt,
                                                     int root, int* __restrict__ prim, float* __restrict__ tout,
                                                     int* __restrict__ count0)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *count0 = npaths;
    const uint32_t seed = *d_seed;
    if (POOL > 0) {
        const int next = (int)(blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 * POOL);
        const int end = npaths < next + 64 * POOL ? npaths : next + 64 * POOL;
        if (next >= end) return;
        const Camera Cr = C; // a register copy: a reference into the kernel argument spills it to scratch
        pool_trace<false>(
            pn, pt, root, next, end,
            [&](int pid, RayT& r, const float4*&, const float4*&, int&) {
                float4 o, d;
                camera_ray(Cr, pid, seed, o, d);
                ray_init(r, o, d);
            },
            [&](bool fin, int pid, const RayT& r, const TravState& S) {
                if (fin) {
                    prim[pid] = S.hit ? S.child : -1;
                    tout[pid] = r.tmax;
                }
            });
        return;
    }
    // work-groups in per-XCD runs (16K paths, a compact patch of pixels)
    const int pid = (int)xcd_chunked_group(blockIdx.x, gridDim.x, XCD_RUN) * (int)blockDim.x + (int)threadIdx.x;
    if (pid >= npaths) return;
    float4 o, d;
    camera_ray(C, pid, seed, o, d);
    RayT r;
    ray_init(r, o, d);
    int st[3] = {0, 0, 0};
    int child = -1;
    const int h = trace_packed<false, false, true>(pn, pt, root, r, child, st);
    prim[pid] = h ? child : -1;
    tout[pid] = r.tmax;
}

// Any hit over the segmented shadow queue with the direct-light add fused
// in: an unoccluded entry adds its contribution to its path's radiance (what
// k_nee_accum does after mt_trace_any_seg).  queue == nullptr: slot = pid.
// With the light-grid guided environment (pend != nullptr), an unoccluded
// environment sample also adds its luminance to its grid entry (s_lg: entry,
// luminance bits; the commented-out accumLightGrid, Render.cl:458-459).
// POOL 0: one entry per lane; POOL > 0: each wave traces a run of 64 * POOL
// entries of one segment, refilling finished lanes (pool_trace).
template <int POOL>
__global__ __launch_bounds__(RT) void k_trace_shadow(const float4* __restrict__ pn, const float4* __restrict__ pt,
                                                     int root, const int* __restrict__ seg_counts, int seg_cap,
                                                     const float4* __restrict__ s_o, const float4* __restrict__ s_d,
                                                     const float4* __restrict__ contrib, const int* __restrict__ queue,
                                                     float4* __restrict__ L, const int2* __restrict__ s_lg,
                                                     unsigned long long* __restrict__ pend)
{
    if (POOL > 0) {
        int next, end;
        pool_seg_run(seg_counts, SHADOW_SEGS, seg_cap, 64 * POOL, next, end);
        if (next >= end) return; // wave-uniform
        pool_trace<true>(
            pn, pt, root, next, end,
            [&](int k, RayT& r, const float4*&, const float4*&, int&) { ray_init(r, s_o[k], s_d[k]); },
            [&](bool fin, int k, const RayT&, const TravState& S) {
                const bool vis = fin && !S.hit;
                if (vis) {
                    const float4 c = contrib[k];
                    const int slot = __float_as_int(c.w);
                    const int pid = queue ? queue[slot] : slot;
                    const float4 v = L[pid];
                    L[pid] = make_float4(v.x + c.x, v.y + c.y, v.z + c.z, v.w);
                }
                if (pend) { // the whole wave is active here
                    int2 e = make_int2(-1, 0);
                    if (vis) e = s_lg[k];
                    lg_wave_accum(e.x >= 0, e.x, lg_fixed(__int_as_float(e.y)), pend);
                }
            });
        return;
    }
    // segment sweep (SEG_SWEEP groups per segment): one entry per lane per chunk
    const unsigned v = seg_sweep_group();
    const int p = (int)(v % SHADOW_SEGS);
    int cnt = seg_counts[p];
    cnt = cnt < seg_cap ? cnt : seg_cap;
    const int stride = (int)(gridDim.x / SHADOW_SEGS) * (int)blockDim.x;
    for (int off = (int)(v / SHADOW_SEGS) * (int)blockDim.x; off < cnt; off += stride) {
        const int k = p * seg_cap + off + (int)threadIdx.x;
        const bool live = off + (int)threadIdx.x < cnt;
        if (__ballot(live) == 0ull) continue; // wave-uniform
        bool vis = false;
        if (live) {
            RayT r;
            ray_init(r, s_o[k], s_d[k]);
            int st[3] = {0, 0, 0};
            int child = -1;
            vis = !trace_packed<true, false, true>(pn, pt, root, r, child, st);
            if (vis) {
                const float4 c = contrib[k];
                const int slot = __float_as_int(c.w);
                const int pid = queue ? queue[slot] : slot;
                const float4 v = L[pid];
                L[pid] = make_float4(v.x + c.x, v.y + c.y, v.z + c.z, v.w);
            }
        }
        if (pend) { // the whole wave is active here
            int2 e = make_int2(-1, 0);
            if (vis) e = s_lg[k];
            lg_wave_accum(e.x >= 0, e.x, lg_fixed(__int_as_float(e.y)), pend);
        }
    }
}

struct ShadeArgs {
    SceneView V;
    PathState S;
    Camera C;             // first: the camera (bounce 0 of the fused loop)
    int first;            // 1: slot = pid, ray = camera ray, throughput 1, radiance not yet stored
    const int* queue;
    const int* d_count;
    float4* o_tmin;       // in: this bounce's rays; out: next rays (same slot)
    float4* d_tmax;
    const int* prim;
    const float* tin;
    int* alive;           // out: 1 if the slot's path continues (unfused loop)
    const int* in_seg;    // fused loop, bounces >= 1: segment lengths of this bounce's queue
    int* next_seg;        // fused loop: segment lengths of the next queue, which
    int* q_next;          //   continuing paths append to (path id, next ray)
    float4* o_next;
    float4* d_next;
    float4* s_o_tmin;     // out: segmented queue of shadow rays
    float4* s_d_tmax;
    float4* s_contrib;    // out: unoccluded contribution, w = the ray's slot (int bits)
    int2* s_lg;           // out (guided environment): grid entry, luminance bits of the light sample
    const int* hits;      // fused loop, bounces >= 1: the queue slots that hit (segmented, k_shade_split)
    const int* hit_seg;   //   and the segment lengths; k_shade then shades only those slots
    int* shadow_count;    // SHADOW_SEGS segment lengths
    int seg_cap;
    int bounce;
    int rr_start;
    int flags;
    const uint32_t* seed; // frame seed (device)
};

// One slot of the fused shading stage; its shadow ray, if any, goes to the
// next free entry of its work-group's shadow-queue segment.
__device__ __forceinline__ void shade_slot(const ShadeArgs& A, uint32_t seed, int slot, bool live)
{
    PathOut P;
    P.has_shadow = false;
    P.alive = false;
    P.acc_gidx = -1;
    P.acc_q = 0ull;
    int pid = -1;
    if (live) {
        float4 ro, rd, T4;
        if (A.first) { // what k_camera would have stored: InitPathData, Initialize.cl:25-48
            pid = slot;
            const Camera C = A.C; // a register copy: a reference into the kernel argument spills it to scratch
            camera_ray(C, pid, seed, ro, rd);
            T4 = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        } else {
            pid = A.queue[slot];
            ro = A.o_tmin[slot];
            rd = A.d_tmax[slot];
            T4 = A.S.T[pid];
        }
        shade_path(A.V, seed, A.rr_start, A.flags, pid, A.bounce, ro, rd, T4, A.prim[slot], A.tin[slot], P);
        if (P.alive) {
            A.S.T[pid] = P.T;
            if (!A.next_seg) {
                A.o_tmin[slot] = P.ro;
                A.d_tmax[slot] = P.rd;
            }
        }
        if (A.first) { // radiance starts at 0 (the same sums as k_camera's zero plus Ladd)
            A.S.L[pid] = make_float4(0.0f + P.Ladd.x, 0.0f + P.Ladd.y, 0.0f + P.Ladd.z, 0.0f);
        } else if (P.Ladd.x != 0.0f || P.Ladd.y != 0.0f || P.Ladd.z != 0.0f) {
            float4 L = A.S.L[pid];
            A.S.L[pid] = make_float4(L.x + P.Ladd.x, L.y + P.Ladd.y, L.z + P.Ladd.z, L.w);
        }
        if (!A.next_seg) A.alive[slot] = P.alive ? 1 : 0;
    }
    // escaped rays of the guided environment mode add to the light grid
    // (accumLightGrid in getRadianceEnvironmentLight1, Light.cl:264); the
    // whole wave is active here
    if (A.V.G.mode) lg_wave_accum(P.acc_gidx >= 0, P.acc_gidx, P.acc_q, A.V.G.pend);
    const int seg = blockIdx.x % SHADOW_SEGS;
    if (A.next_seg) {
        // the next queue, segmented like the shadow queue: no result depends
        // on its order (everything is indexed by path id), so continuing
        // paths are appended here instead of compacted in a separate pass
        const int kn = seg * A.seg_cap + wave_append(P.alive, A.next_seg + seg);
        if (P.alive) {
            A.q_next[kn] = pid;
            A.o_next[kn] = P.ro;
            A.d_next[kn] = P.rd;
        }
    }
    const int k = seg * A.seg_cap + wave_append(P.has_shadow, A.shadow_count + seg);
    if (P.has_shadow) {
        A.s_o_tmin[k] = P.so;
        A.s_d_tmax[k] = P.sd;
        A.s_contrib[k] = make_float4(P.contrib.x, P.contrib.y, P.contrib.z, __int_as_float(slot));
        if (A.s_lg) A.s_lg[k] = make_int2(P.shadow_gidx, __float_as_int(P.shadow_lum));
    }
}

// SWEEP false (bounce 0, the unfused loop): one slot per lane over a
// capacity-sized grid (the segment capacities assume exactly that;
// work-groups past *d_count leave at once).  SWEEP true (fused bounces >=
// 1): a segment sweep over the queue or over the slots k_shade_split found
// to hit.  Two instantiations, so the sweep's loop does not change the
// register allocation of the bounce-0 kernel.
template <bool SWEEP>
__global__ __launch_bounds__(RT) void k_shade(ShadeArgs A)
{
    int slot;
    bool live;
    if (SWEEP) { // segment sweep over the hitting slots, or over the queue
        const uint32_t seed = *A.seed;
        const int p = blockIdx.x % SHADOW_SEGS;
        int cnt = A.hits ? A.hit_seg[p] : A.in_seg[p];
        cnt = cnt < A.seg_cap ? cnt : A.seg_cap;
        const int stride = (int)(gridDim.x / SHADOW_SEGS) * (int)blockDim.x;
        for (int off = (int)(blockIdx.x / SHADOW_SEGS) * (int)blockDim.x; off < cnt; off += stride) {
            live = off + (int)threadIdx.x < cnt;
            const int k = p * A.seg_cap + off + (int)threadIdx.x;
            slot = A.hits ? (live ? A.hits[k] : 0) : k;
            shade_slot(A, seed, slot, live);
        }
        return;
    } else {
        const int base = blockIdx.x * blockDim.x;
        const int count = *A.d_count;
        if (base >= count) return;
        slot = base + threadIdx.x;
        live = slot < count;
    }
    shade_slot(A, *A.seed, slot, live);
}

// Bounces >= 1 of the fused loop, before k_shade: escaped paths gather the
// environment's radiance here (shade_escape, the shading of a miss) and the
// slots that hit a triangle are appended, per work-group segment, to a dense
// list that k_shade then shades.  Most paths escape after the first bounce
// and the hits are scattered through the queue, so without the split nearly
// every wave would run the whole hit shading for a few lanes; each path's
// shading is the same code either way.
__global__ __launch_bounds__(RT) void k_shade_split(ShadeArgs A, int* __restrict__ hits, int* __restrict__ hit_seg)
{
    const int p = blockIdx.x % SHADOW_SEGS; // segment sweep
    int cnt = A.in_seg[p];
    cnt = cnt < A.seg_cap ? cnt : A.seg_cap;
    const int stride = (int)(gridDim.x / SHADOW_SEGS) * (int)blockDim.x;
    for (int off = (int)(blockIdx.x / SHADOW_SEGS) * (int)blockDim.x; off < cnt; off += stride) {
        const int slot = p * A.seg_cap + off + (int)threadIdx.x;
        const bool live = off + (int)threadIdx.x < cnt;
        const int prim = live ? A.prim[slot] : -1;
        int gidx = -1;
        unsigned long long q = 0ull;
        if (live && prim < 0) {
            const int pid = A.queue[slot];
            v3 Ladd;
            shade_escape(A.V, A.bounce, A.o_tmin[slot], A.d_tmax[slot], A.S.T[pid], Ladd, gidx, q);
            if (Ladd.x != 0.0f || Ladd.y != 0.0f || Ladd.z != 0.0f) {
                const float4 L = A.S.L[pid];
                A.S.L[pid] = make_float4(L.x + Ladd.x, L.y + Ladd.y, L.z + Ladd.z, L.w);
            }
        }
        if (A.V.G.mode) lg_wave_accum(gidx >= 0, gidx, q, A.V.G.pend); // the whole wave is active here
        const bool hit = live && prim >= 0;
        const int k = p * A.seg_cap + wave_append(hit, hit_seg + p);
        if (hit) hits[k] = slot;
    }
}

// Bounces tail..nb-1 of the fused loop in one launch (MT_RENDER_TAIL_FROM):
// each lane takes a path of bounce `tail`'s segmented queue and carries it
// through all of its remaining bounces — closest hit, shading (shade_path),
// shadow ray and its direct-light add — then takes its segment's next path.
// Every path goes through the same operations in the same order as in the
// per-bounce kernels (its radiance gains the shading's Ladd, then the
// unoccluded shadow contribution, bounce after bounce; its random numbers
// are keyed by (seed, path id, bounce); the light-grid sums are integers),
// so the image and the grid are bit-identical.  What changes is the time
// of the sparse late bounces: one launch whose length is the slowest path's
// chain of dependent loads, instead of three launches per bounce, each as
// long as that bounce's slowest ray.  Lanes of a wave stay in step (every
// iteration is trace, shade, shadow for the lanes holding a path), only
// their bounce numbers differ.  live[b] / shadow[b] count the paths
// entering bounce b > tail and the shadow rays of bounce b >= tail (the
// per-bounce kernels' queue lengths), summed in LDS, one atomic per
// work-group and bounce.
constexpr int TAIL_SWEEP = 8; // work-groups per segment
__global__ __launch_bounds__(RT) void k_tail(ShadeArgs A, const float4* __restrict__ pn,
                                             const float4* __restrict__ pt, int root, int nb,
                                             int* __restrict__ live_out, int* __restrict__ shadow_out)
{
    __shared__ int s_live[64], s_shadow[64];
    if (threadIdx.x < 64) {
        s_live[threadIdx.x] = 0;
        s_shadow[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t seed = *A.seed;
    const int p = blockIdx.x % SHADOW_SEGS;
    int cnt = A.in_seg[p];
    cnt = cnt < A.seg_cap ? cnt : A.seg_cap;
    const int stride = (int)(gridDim.x / SHADOW_SEGS) * (int)blockDim.x;
    int e = (int)(blockIdx.x / SHADOW_SEGS) * (int)blockDim.x + (int)threadIdx.x; // this lane's next entry
    bool have = false;
    int pid = 0, b = 0;
    float4 ro = make_float4(0, 0, 0, 0), rd = ro, T4 = ro, L = ro;
    for (;;) {
        if (!have && e < cnt) { // refill: the next path of this lane's share of the segment
            const int slot = p * A.seg_cap + e;
            pid = A.queue[slot];
            ro = A.o_tmin[slot];
            rd = A.d_tmax[slot];
            T4 = A.S.T[pid];
            L = A.S.L[pid];
            b = A.bounce;
            have = true;
            e += stride;
        }
        if (__ballot(have) == 0ull) break; // wave-uniform
        int prim = -1;
        float t = 0.0f;
        if (have) { // closest hit, as k_trace_pool / k_trace_camera find it
            RayT r;
            ray_init(r, ro, rd);
            int st[3] = {0, 0, 0};
            int child = -1;
            prim = trace_packed<false, false, true>(pn, pt, root, r, child, st) ? child : -1;
            t = r.tmax;
        }
        PathOut P;
        P.has_shadow = false;
        P.alive = false;
        P.acc_gidx = -1;
        P.acc_q = 0ull;
        if (have) {
            shade_path(A.V, seed, A.rr_start, A.flags, pid, b, ro, rd, T4, prim, t, P);
            if (P.Ladd.x != 0.0f || P.Ladd.y != 0.0f || P.Ladd.z != 0.0f)
                L = make_float4(L.x + P.Ladd.x, L.y + P.Ladd.y, L.z + P.Ladd.z, L.w);
        }
        if (A.V.G.mode) lg_wave_accum(P.acc_gidx >= 0, P.acc_gidx, P.acc_q, A.V.G.pend); // whole wave active
        bool vis = false;
        if (P.has_shadow) { // any hit; an unoccluded sample adds its contribution
            atomicAdd(&s_shadow[b], 1);
            RayT r;
            ray_init(r, P.so, P.sd);
            int st[3] = {0, 0, 0};
            int child = -1;
            vis = !trace_packed<true, false, true>(pn, pt, root, r, child, st);
            if (vis) L = make_float4(L.x + P.contrib.x, L.y + P.contrib.y, L.z + P.contrib.z, L.w);
        }
        if (A.V.G.mode && A.V.G.pend)
            lg_wave_accum(vis && P.shadow_gidx >= 0, P.shadow_gidx, lg_fixed(P.shadow_lum), A.V.G.pend);
        if (have) {
            if (P.alive && b + 1 < nb) { // the path's next bounce, in this lane
                ro = P.ro;
                rd = P.rd;
                T4 = P.T;
                b++;
                atomicAdd(&s_live[b], 1);
            } else { // the path ends: its radiance goes back (throughput is not read again)
                A.S.L[pid] = L;
                have = false;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && (int)threadIdx.x < nb) {
        if (s_live[threadIdx.x]) atomicAdd(live_out + threadIdx.x, s_live[threadIdx.x]);
        if (s_shadow[threadIdx.x]) atomicAdd(shadow_out + threadIdx.x, s_shadow[threadIdx.x]);
    }
}

// Unoccluded shadow-queue entries add their contribution to their path.
// (pend != nullptr: an unoccluded environment sample of the guided mode also
// adds its luminance to its grid entry, as in k_trace_shadow.)
__global__ __launch_bounds__(RT) void k_nee_accum(const int* __restrict__ queue, const int* __restrict__ seg_counts,
                                                  int seg_cap, const float4* __restrict__ contrib,
                                                  const int* __restrict__ occluded, PathState S,
                                                  const int2* __restrict__ s_lg, unsigned long long* __restrict__ pend)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool vis = seg_live(seg_counts, seg_cap, k) && !occluded[k];
    if (vis) {
        const float4 c = contrib[k];
        const int pid = queue[__float_as_int(c.w)];
        float4 L = S.L[pid];
        S.L[pid] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
    }
    if (pend) {
        int2 e = make_int2(-1, 0);
        if (vis) e = s_lg[k];
        lg_wave_accum(e.x >= 0, e.x, lg_fixed(__int_as_float(e.y)), pend);
    }
}

__global__ void k_zero_i32(int* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Totals of the traversal counters live in ILP_SLOTS rows of 8 words
// ({I, L, P} of closest hit, then of any hit); work-group b adds into row
// b % ILP_SLOTS, so the waves' atomics do not all meet on three addresses;
// the host sums the rows.
constexpr int ILP_SLOTS = 64;

// Sums per-entry traversal counters {I, L, P} of the live entries: i <
// *d_count, or (seg_counts set) the live entries of a segmented queue.
__global__ __launch_bounds__(RT) void k_sum_ilp(const int* __restrict__ st, const int* __restrict__ d_count,
                                                const int* __restrict__ seg_counts, int seg_cap,
                                                unsigned long long* __restrict__ tot_rows)
{
    unsigned long long* tot = tot_rows + 8 * (blockIdx.x % ILP_SLOTS);
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

// Wavefront state of one render job: one device allocation, laid out once.
// Offsets are bytes into `base`.
struct RenderPlan {
    SceneView V;
    Camera C;
    mt_render_desc desc;
    const mt_scene* scene;
    float* image;
    int np, nb;
    int64_t npix;
    int seg_cap; // shadow-queue segment capacity (SHADOW_SEGS segments)
    size_t sq;   // shadow-queue capacity
    bool want_ilp, profile;
    char* base;
    size_t bytes;
    size_t oT, oL, oA_o, oA_d, oB_o, oB_d, oQA, oQB, oPrim, oT_, oAlive, oOcc, oSo, oSd, oSc, oCnt, oShc, oQsc,
        oIlp, oTot, oSeed, oSlg, oHit, oHsc, oTail;
    mt_lightgrid* grid; // guided environment mode: the grid the render samples and accumulates into
    bool fused; // camera rays traced in place, shadow adds fused, next queue appended by the shading stage
    bool pool;  // fused kernels trace wave runs with lane refill (pool_trace)
    int tail;   // fused loop: bounces tail..nb-1 run in one k_tail launch (0: none)
    bool borrowed; // base is the context's reusable workspace (mt_render), not owned
    hipEvent_t ev[4 * 64];
    int nev;
    // fused loop without per-stage timing: bounce b's any hit runs on the
    // context's side stream, concurrently with bounce b+1's closest hit
    // (the context's fork[b] after b's shading, join[b] before b+1's)
    hipEvent_t* fork;
    hipEvent_t* join;
    hipStream_t side;
};

int32_t plan_render(const mt_scene* scene, const float* materials, const mt_envmap* env,
                    const int32_t* light_faces, const int32_t* d_num_lights, const mt_render_desc* desc,
                    float* image, RenderPlan& R)
{
    MT_ARG_CHECK(scene && materials && desc && image, "mt_render: NULL argument");
    MT_ARG_CHECK(desc->width > 0 && desc->height > 0 && desc->spp > 0, "mt_render: empty image");
    MT_ARG_CHECK(desc->max_bounces >= 1 && desc->max_bounces <= 63, "mt_render: max_bounces %d not in [1,63]",
                 desc->max_bounces);
    MT_ARG_CHECK(!light_faces == !d_num_lights, "mt_render: light_faces and d_num_lights go together");
    R.npix = (int64_t)desc->width * desc->height;
    const int64_t np64 = R.npix * desc->spp;
    MT_ARG_CHECK(np64 <= (1ll << 30), "mt_render: %lld paths exceed the 2^30 wavefront limit", (long long)np64);
    R.np = (int)np64;
    R.nb = desc->max_bounces;
    R.desc = *desc;
    R.scene = scene;
    R.image = image;
    R.base = nullptr;
    R.borrowed = false;
    R.nev = 0;
    R.fork = R.join = nullptr;
    R.side = nullptr;
    int32_t rc = camera_from_desc(desc, R.C);
    if (rc) return rc;
    R.V.points = reinterpret_cast<const float4*>(scene->desc.points);
    R.V.normals = reinterpret_cast<const float4*>(scene->desc.normals);
    R.V.faces = scene->desc.faces;
    R.V.mats = reinterpret_cast<const MtMaterial*>(materials);
    R.V.light_faces = light_faces;
    R.V.num_lights = d_num_lights;
    R.V.env = env_view(env);
    R.grid = nullptr;
    R.V.G = LgView{nullptr, nullptr, nullptr, 0.0f, 0.0f, 0.0f, nullptr, 0};
    const bool guided = (desc->flags & MT_RENDER_LIGHTGRID_ENV) != 0;
    MT_ARG_CHECK(guided || !(desc->flags & MT_RENDER_LIGHTGRID_TILE_PDF),
                 "mt_render: MT_RENDER_LIGHTGRID_TILE_PDF needs MT_RENDER_LIGHTGRID_ENV");
    if (guided) {
        MT_ARG_CHECK(env && env->grid, "mt_render: MT_RENDER_LIGHTGRID_ENV needs an environment with a light grid "
                                       "(mt_env_attach_lightgrid)");
        mt_lightgrid* g = env->grid;
        R.grid = g;
        R.V.G = LgView{g->func, g->sat, g->tab_valid ? g->tab : nullptr, g->cam[0], g->cam[1], g->cam[2], g->pend,
                       (desc->flags & MT_RENDER_LIGHTGRID_TILE_PDF) ? 2 : 1};
    }

    const size_t n = (size_t)R.np;
    size_t bytes = 0;
    auto take = [&](size_t b) { size_t off = bytes; bytes += (b + 255) & ~(size_t)255; return off; };
    // shadow rays, and the path queues of bounces >= 1 in the fused loop:
    // segmented queues filled by k_shade's work-groups (segment = group mod
    // SHADOW_SEGS, capacity = the slots of the groups that map to it)
    R.seg_cap = shadow_seg_cap(R.np, RT);
    R.sq = (size_t)SHADOW_SEGS * R.seg_cap;
    const size_t nq = R.sq > n ? R.sq : n;
    R.oT = take(16 * n);
    R.oL = take(16 * n);
    R.oA_o = take(16 * nq);
    R.oA_d = take(16 * nq);
    R.oB_o = take(16 * nq);
    R.oB_d = take(16 * nq);
    R.oQA = take(4 * nq);
    R.oQB = take(4 * nq);
    R.oPrim = take(4 * nq);
    R.oT_ = take(4 * nq);
    R.oAlive = take(4 * n);
    R.oOcc = take(4 * R.sq);
    R.oSo = take(16 * R.sq);
    R.oSd = take(16 * R.sq);
    R.oSc = take(16 * R.sq);
    R.oCnt = take(4 * 64);
    R.oShc = take(4 * 64 * (size_t)SHADOW_SEGS);
    R.oQsc = take(4 * 64 * (size_t)SHADOW_SEGS);
    R.oHit = take(4 * R.sq);
    R.oHsc = take(4 * 64 * (size_t)SHADOW_SEGS);
    R.oTail = take(4 * 2 * 64); // k_tail's live paths per bounce, then its shadow rays per bounce
    R.want_ilp = (desc->flags & MT_RENDER_TRAVERSAL_STATS) != 0;
    R.fused = (desc->mode == MT_TRACE_PACKED || desc->mode == MT_TRACE_PACKED_STACK ||
               desc->mode == MT_TRACE_PACKED_SORTED || desc->mode == MT_TRACE_PACKED_POOL) && !R.want_ilp;
    R.pool = desc->mode == MT_TRACE_PACKED_POOL;
    R.tail = (desc->flags >> MT_RENDER_TAIL_SHIFT) & 63;
    if (!R.fused || R.tail >= R.nb) R.tail = 0; // the tail applies to the fused loop's bounces >= 1
    R.profile = (desc->flags & MT_RENDER_PROFILE) != 0;
    R.oIlp = R.want_ilp ? take(12 * (R.sq > n ? R.sq : n)) : 0;
    R.oTot = take(8 * 8 * ILP_SLOTS);
    R.oSeed = take(4);
    R.oSlg = guided ? take(8 * R.sq) : 0;
    R.bytes = bytes;
    return MT_OK;
}

void free_render(RenderPlan& R)
{
    for (int k = 0; k < R.nev; k++) (void)hipEventDestroy(R.ev[k]);
    R.nev = 0;
    R.fork = R.join = nullptr; // the context's
    if (R.base && !R.borrowed) (void)hipFree(R.base);
    R.base = nullptr;
}

// borrow: the state lives in the context's grow-only workspace (calls on a
// context are serialised and mt_render synchronises before it returns), so
// repeated renders allocate nothing; otherwise it is owned.
int32_t alloc_render(RenderPlan& R, mt_context* ctx, bool borrow)
{
    R.borrowed = borrow;
    if (borrow) {
        void* p = nullptr;
        int32_t rc = mt_ctx_temp(ctx, R.bytes, &p);
        if (rc) {
            R.base = nullptr;
            return rc;
        }
        R.base = reinterpret_cast<char*>(p);
    } else if (hipMalloc(&R.base, R.bytes) != hipSuccess) {
        R.base = nullptr;
        mt_set_error("mt_render: cannot allocate %.2f GB of wavefront state", R.bytes / 1e9);
        return MT_ERR_ALLOC;
    }
    if (R.profile) {
        for (; R.nev < 4 * R.nb; R.nev++)
            if (hipEventCreate(&R.ev[R.nev]) != hipSuccess) {
                free_render(R);
                mt_set_error("mt_render: hipEventCreate failed");
                return MT_ERR_HIP;
            }
    }
    R.side = nullptr;
    if (R.fused && !R.profile && R.nb > 1) {
        int32_t rc = mt_ctx_side_stream(ctx, &R.side);
        if (rc) {
            R.side = nullptr;
            free_render(R);
            return rc;
        }
        R.fork = ctx->fork;
        R.join = ctx->join;
    }
    return MT_OK;
}

uint32_t* seed_ptr(const RenderPlan& R) { return reinterpret_cast<uint32_t*>(R.base + R.oSeed); }

// Enqueues the whole job on ctx->stream: kernel launches only (no
// allocation, no synchronisation, no memset), so it can be captured.  The
// frame seed is read from seed_ptr(R).
int32_t enqueue_render(mt_context* ctx, RenderPlan& R)
{
    hipStream_t st = ctx->stream;
    char* base = R.base;
    const mt_render_desc* desc = &R.desc;
    const mt_scene* scene = R.scene;
    const int np = R.np, nb = R.nb, seg_cap = R.seg_cap;
    auto F4 = [&](size_t off) { return reinterpret_cast<float4*>(base + off); };
    auto I = [&](size_t off) { return reinterpret_cast<int*>(base + off); };
    PathState S{F4(R.oT), F4(R.oL)};
    float4* ro[2] = {F4(R.oA_o), F4(R.oB_o)};
    float4* rd[2] = {F4(R.oA_d), F4(R.oB_d)};
    int* q[2] = {I(R.oQA), I(R.oQB)};
    int* counts = I(R.oCnt);
    int* shc = I(R.oShc);
    int* qsc = I(R.oQsc); // row b: segment lengths of bounce b's queue (fused loop, b >= 1)
    int* hits = I(R.oHit); // bounce b's hitting slots (fused loop, b >= 1), segment lengths in row b of hsc
    int* hsc = I(R.oHsc);
    auto chk = [&](void) -> int32_t {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            mt_set_error("mt_render: launch failed: %s", hipGetErrorString(e));
            return MT_ERR_HIP;
        }
        return MT_OK;
    };
    const int grid = mt_blocks(np, RT);
    const int gshade = grid;              // one slot per lane; empty work-groups retire at once
    const int gshadow = mt_blocks((int64_t)R.sq, RT);
    // the fused loop's segment sweeps: SEG_SWEEP work-groups per segment, or
    // fewer when a segment holds fewer chunks of RT slots (small jobs: the
    // extra groups could never find work, only be dispatched)
    const int seg_chunks = (int)(((int64_t)seg_cap + RT - 1) / RT);
    const int gsweep = SHADOW_SEGS * (seg_chunks < SEG_SWEEP ? seg_chunks : SEG_SWEEP);
    const int gtail = SHADOW_SEGS * (seg_chunks < TAIL_SWEEP ? seg_chunks : TAIL_SWEEP);
    // wave-pool kernels: runs of 64 * POOL_RAYS entries, 4 per work-group
    const int64_t pool_group = 4 * 64 * (int64_t)POOL_RAYS;
    const int gcam_pool = mt_blocks(np, (int)pool_group);
    const int gshadow_pool = SHADOW_SEGS * (int)((seg_cap + pool_group - 1) / pool_group);
    unsigned long long* tot = reinterpret_cast<unsigned long long*>(base + R.oTot);
    int* ilp = R.want_ilp ? I(R.oIlp) : nullptr;
    hipEvent_t* ev = R.ev;
    const bool profile = R.profile;
    auto mark = [&](hipEvent_t e) -> int32_t {
        const hipError_t x = hipEventRecord(e, st);
        if (x == hipSuccess) return MT_OK;
        mt_set_error("mt_render: hipEventRecord failed: %s", hipGetErrorString(x));
        return MT_ERR_HIP;
    };

    // bounce b's any hit on the side stream while bounce b+1's closest hit
    // runs here: they read and write disjoint state (the shadow queue and
    // the radiance / the next queue and the hit records); b+1's shading
    // waits for it (its radiance adds come after b's, and it refills the
    // shadow queue)
    const bool overlap = R.side != nullptr && R.fork != nullptr;
    int pending = -1; // bounce whose any hit runs on the side stream
    auto join_pending = [&]() -> int32_t {
        if (pending < 0) return MT_OK;
        const hipError_t x = hipStreamWaitEvent(st, R.join[pending], 0);
        pending = -1;
        if (x == hipSuccess) return MT_OK;
        mt_set_error("mt_render: hipStreamWaitEvent failed: %s", hipGetErrorString(x));
        return MT_ERR_HIP;
    };

    k_zero_i32<<<1, 64, 0, st>>>(counts, 64);
    k_zero_i32<<<mt_blocks(64 * SHADOW_SEGS, RT), RT, 0, st>>>(shc, 64 * SHADOW_SEGS);
    k_zero_i32<<<mt_blocks(64 * SHADOW_SEGS, RT), RT, 0, st>>>(qsc, 64 * SHADOW_SEGS);
    k_zero_i32<<<mt_blocks(64 * SHADOW_SEGS, RT), RT, 0, st>>>(hsc, 64 * SHADOW_SEGS);
    k_zero_i32<<<1, 128, 0, st>>>(I(R.oTail), 128);
    k_zero_i32<<<mt_blocks(16 * ILP_SLOTS, RT), RT, 0, st>>>(reinterpret_cast<int*>(tot), 16 * ILP_SLOTS);
    // Fused loop (default traversal, no traversal counters): bounce 0 traces
    // camera rays generated in place and shades with slot = pid; shadow rays
    // add their light inside the any-hit kernel; continuing paths append
    // themselves to the next bounce's segmented queue.  Otherwise camera rays
    // are stored first, every stage is a separate launch and each bounce's
    // survivors are compacted in slot order.
    const bool fused = R.fused;
    int32_t rc = MT_OK;
    if (!fused) {
        k_camera<<<grid, RT, 0, st>>>(R.C, np, seed_ptr(R), ro[0], rd[0], q[0], S, counts);
        rc = chk();
    }
    int cur = 0; // parity of the queue buffers holding the current bounce
    // the shading stage's arguments for bounce b (queue buffers of parity cur)
    auto shade_args = [&](int b, bool first) {
        ShadeArgs A;
        A.V = R.V;
        A.S = S;
        A.C = R.C;
        A.first = first ? 1 : 0;
        A.queue = q[cur];
        A.d_count = counts + b;
        A.o_tmin = ro[cur];
        A.d_tmax = rd[cur];
        A.prim = I(R.oPrim);
        A.tin = reinterpret_cast<const float*>(base + R.oT_);
        A.alive = I(R.oAlive);
        A.in_seg = fused && !first ? qsc + (size_t)b * SHADOW_SEGS : nullptr;
        A.next_seg = fused ? qsc + (size_t)(b + 1) * SHADOW_SEGS : nullptr;
        A.q_next = q[1 - cur];
        A.o_next = ro[1 - cur];
        A.d_next = rd[1 - cur];
        A.s_o_tmin = F4(R.oSo);
        A.s_d_tmax = F4(R.oSd);
        A.s_contrib = F4(R.oSc);
        A.s_lg = R.grid ? reinterpret_cast<int2*>(base + R.oSlg) : nullptr;
        A.shadow_count = shc + (size_t)b * SHADOW_SEGS;
        A.seg_cap = seg_cap;
        A.bounce = b;
        A.rr_start = desc->rr_start;
        A.flags = desc->flags;
        A.seed = seed_ptr(R);
        A.hits = nullptr;
        A.hit_seg = nullptr;
        return A;
    };
    for (int b = 0; b < nb && rc == MT_OK; b++) {
        int* dcount = counts + b;
        const bool first = fused && b == 0;
        if (R.tail && b == R.tail) { // every remaining bounce in one launch
            if ((rc = join_pending())) break;
            k_tail<<<gtail, RT, 0, st>>>(shade_args(b, false), scene->pnodes, scene->ptris,
                                                          scene->root_ref, nb, I(R.oTail), I(R.oTail) + 64);
            rc = chk();
            break;
        }
        if (profile && (rc = mark(ev[4 * b]))) break;
        if (first) {
            if (R.pool)
                k_trace_camera<POOL_RAYS><<<gcam_pool, RT, 0, st>>>(R.C, np, seed_ptr(R), scene->pnodes, scene->ptris,
                                                                    scene->root_ref, I(R.oPrim),
                                                                    reinterpret_cast<float*>(base + R.oT_), counts);
            else
                k_trace_camera<0><<<grid, RT, 0, st>>>(R.C, np, seed_ptr(R), scene->pnodes, scene->ptris,
                                                       scene->root_ref, I(R.oPrim),
                                                       reinterpret_cast<float*>(base + R.oT_), counts);
            rc = chk();
        } else if (fused) {
            rc = mt_trace_closest_seg(ctx, scene, ro[cur], rd[cur], MtSegQueue{qsc + (size_t)b * SHADOW_SEGS,
                                                                               SHADOW_SEGS, seg_cap},
                                      I(R.oPrim), reinterpret_cast<float*>(base + R.oT_), desc->mode);
        } else {
            rc = mt_trace_closest(ctx, scene, reinterpret_cast<float*>(ro[cur]), reinterpret_cast<float*>(rd[cur]),
                                  dcount, np, I(R.oPrim), reinterpret_cast<float*>(base + R.oT_), ilp, desc->mode);
        }
        if (rc) break;
        if (profile && (rc = mark(ev[4 * b + 1]))) break;
        if (ilp) k_sum_ilp<<<grid, RT, 0, st>>>(ilp, dcount, nullptr, 0, tot);
        if ((rc = join_pending())) break;
        ShadeArgs A = shade_args(b, first);
        int* seg_counts = A.shadow_count;
        if (A.in_seg) { // escaped paths here, the hits densely in k_shade
            int* hseg = hsc + (size_t)b * SHADOW_SEGS;
            k_shade_split<<<gsweep, RT, 0, st>>>(A, hits, hseg);
            if ((rc = chk())) break;
            A.hits = hits;
            A.hit_seg = hseg;
        }
        if (A.in_seg) k_shade<true><<<gsweep, RT, 0, st>>>(A);
        else k_shade<false><<<gshade, RT, 0, st>>>(A);
        if ((rc = chk())) break;
        if (profile && (rc = mark(ev[4 * b + 2]))) break;
        if (fused) {
            hipStream_t ss = st;
            if (overlap) { // fork: the any hit follows this shading on the side stream
                hipError_t x = hipEventRecord(R.fork[b], st);
                if (x == hipSuccess) x = hipStreamWaitEvent(R.side, R.fork[b], 0);
                if (x != hipSuccess) {
                    mt_set_error("mt_render: fork to the side stream failed: %s", hipGetErrorString(x));
                    rc = MT_ERR_HIP;
                    break;
                }
                ss = R.side;
            }
            if (R.pool)
                k_trace_shadow<POOL_RAYS><<<gshadow_pool, RT, 0, ss>>>(
                    scene->pnodes, scene->ptris, scene->root_ref, seg_counts, seg_cap, F4(R.oSo), F4(R.oSd),
                    F4(R.oSc), first ? nullptr : q[cur], S.L, A.s_lg, R.grid ? R.grid->pend : nullptr);
            else
                k_trace_shadow<0><<<gsweep, RT, 0, ss>>>(
                    scene->pnodes, scene->ptris, scene->root_ref, seg_counts, seg_cap, F4(R.oSo), F4(R.oSd),
                    F4(R.oSc), first ? nullptr : q[cur], S.L, A.s_lg, R.grid ? R.grid->pend : nullptr);
            if ((rc = chk())) break;
            if (overlap) {
                const hipError_t x = hipEventRecord(R.join[b], R.side);
                if (x != hipSuccess) {
                    mt_set_error("mt_render: join event failed: %s", hipGetErrorString(x));
                    rc = MT_ERR_HIP;
                    break;
                }
                pending = b;
            }
            if (profile && (rc = mark(ev[4 * b + 3]))) break;
        } else {
            // any hit over the segmented shadow queue (segment lengths on the device)
            rc = mt_trace_any_seg(ctx, scene, F4(R.oSo), F4(R.oSd), MtSegQueue{seg_counts, SHADOW_SEGS, seg_cap},
                                  I(R.oOcc), ilp, desc->mode);
            if (rc) break;
            if (profile && (rc = mark(ev[4 * b + 3]))) break;
            if (ilp) k_sum_ilp<<<gshadow, RT, 0, st>>>(ilp, nullptr, seg_counts, seg_cap, tot + 3);
            k_nee_accum<<<gshadow, RT, 0, st>>>(q[cur], seg_counts, seg_cap, F4(R.oSc), I(R.oOcc), S, A.s_lg,
                                                R.grid ? R.grid->pend : nullptr);
            if ((rc = chk())) break;
        }
        if (fused) {
            cur = 1 - cur;
        } else if (b + 1 < nb) { // unfused loop: stable compaction of the survivors
            mt_payload P[3] = {{q[cur], q[1 - cur], 4}, {ro[cur], ro[1 - cur], 16}, {rd[cur], rd[1 - cur], 16}};
            rc = mt_compact(ctx, I(R.oAlive), 0, dcount, np, nullptr, counts + b + 1, P, 3);
            if (rc) break;
            cur = 1 - cur;
        }
    }
    { // the last any hit joins back (also after an error: nothing may still run on the side stream)
        const int32_t jr = join_pending();
        if (rc == MT_OK) rc = jr;
    }
    if (rc == MT_OK) {
        k_resolve<<<mt_blocks(R.npix, resolve_pixels_per_group(desc->spp)), RT, 0, st>>>(S.L, desc->width, desc->height, desc->spp,
                                                         reinterpret_cast<float4*>(R.image));
        rc = chk();
    }
    // the frame's light-grid sums join the grid's float accumulation
    if (rc == MT_OK && R.grid) rc = mt_lightgrid_fold(st, R.grid);
    return rc;
}

// Reads the ray counts (and, when enabled, traversal counters and event
// timings) of the last enqueued render; synchronises the stream.
int32_t read_stats(mt_context* ctx, RenderPlan& R, mt_render_stats* stats)
{
    hipStream_t st = ctx->stream;
    const int nb = R.nb;
    int hc[64];
    std::vector<int> hs((size_t)nb * SHADOW_SEGS);
    std::vector<unsigned long long> ht(8 * ILP_SLOTS);
    hipError_t e = hipMemcpyAsync(hc, R.base + R.oCnt, sizeof(hc), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hs.data(), R.base + R.oShc, hs.size() * sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ht.data(), R.base + R.oTot, ht.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        mt_set_error("mt_render: stats readback failed: %s", hipGetErrorString(e));
        return MT_ERR_HIP;
    }
    stats->extension_rays = 0;
    stats->shadow_rays = 0;
    stats->paths = R.np;
    if (R.fused) { // bounces >= 1: the segment lengths of the appended queues
        std::vector<int> hq((size_t)nb * SHADOW_SEGS);
        e = hipMemcpyAsync(hq.data(), R.base + R.oQsc, hq.size() * sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            mt_set_error("mt_render: stats readback failed: %s", hipGetErrorString(e));
            return MT_ERR_HIP;
        }
        for (int b = 1; b < nb; b++) {
            int64_t c = 0;
            for (int p = 0; p < SHADOW_SEGS; p++) c += hq[(size_t)b * SHADOW_SEGS + p];
            hc[b] = (int)c;
        }
    }
    if (R.tail) { // k_tail's bounces: paths entering bounces > tail, shadow rays of bounces >= tail
        int tl[128];
        e = hipMemcpyAsync(tl, R.base + R.oTail, sizeof(tl), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            mt_set_error("mt_render: stats readback failed: %s", hipGetErrorString(e));
            return MT_ERR_HIP;
        }
        for (int b = R.tail + 1; b < nb; b++) hc[b] += tl[b];
        for (int b = R.tail; b < nb; b++) stats->shadow_rays += tl[64 + b];
    }
    for (int b = 0; b < 64; b++) stats->bounce_counts[b] = b < nb ? hc[b] : 0;
    for (int b = 0; b < nb; b++) stats->extension_rays += hc[b];
    for (int v : hs) stats->shadow_rays += v;
    for (int k = 0; k < 3; k++) {
        stats->closest_ilp[k] = stats->any_ilp[k] = 0;
        for (int row = 0; row < ILP_SLOTS; row++) {
            stats->closest_ilp[k] += (int64_t)ht[8 * row + k];
            stats->any_ilp[k] += (int64_t)ht[8 * row + 3 + k];
        }
    }
    stats->closest_launches = stats->any_launches = 0;
    stats->closest_ms = stats->any_ms = 0.0f;
    if (R.profile) {
        for (int b = 0; b < (R.tail ? R.tail : nb); b++) { // k_tail's bounces have no per-stage events
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, R.ev[4 * b], R.ev[4 * b + 1]) == hipSuccess) {
                stats->closest_ms += ms;
                stats->closest_launches++;
            }
            if (hipEventElapsedTime(&ms, R.ev[4 * b + 2], R.ev[4 * b + 3]) == hipSuccess) {
                stats->any_ms += ms;
                stats->any_launches++;
            }
        }
    }
    return MT_OK;
}

} // namespace

extern "C" int32_t mt_render(mt_context* ctx, const mt_scene* scene, const float* materials,
                             const mt_envmap* env, const int32_t* light_faces, const int32_t* d_num_lights,
                             const mt_render_desc* desc, float* image, mt_render_stats* stats)
{
    MT_ARG_CHECK(ctx, "mt_render: NULL context");
    RenderPlan R;
    int32_t rc = plan_render(scene, materials, env, light_faces, d_num_lights, desc, image, R);
    if (rc) return rc;
    rc = alloc_render(R, ctx, true);
    if (rc) return rc;
    k_set_seed<<<1, 64, 0, ctx->stream>>>(seed_ptr(R), desc->seed);
    rc = enqueue_render(ctx, R);
    if (rc == MT_OK && stats) rc = read_stats(ctx, R, stats);
    const hipError_t se = hipStreamSynchronize(ctx->stream);
    free_render(R);
    if (rc == MT_OK && se != hipSuccess) {
        mt_set_error("mt_render: %s", hipGetErrorString(se));
        rc = MT_ERR_HIP;
    }
    if (rc == MT_OK) rc = mt_scan_check(ctx);
    return rc;
}

// ---------------------------------------------------------------- captured render
struct mt_render_graph {
    mt_context* ctx;
    RenderPlan R;
    void* scratch; // the graph's own look-back workspace (status words, ticket, epoch)
    size_t scratch_bytes;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" int32_t mt_render_graph_destroy(mt_render_graph* g)
{
    if (!g) return MT_OK;
    MtFirstError fe;
    fe(hipStreamSynchronize(g->ctx->stream));
    if (g->exec) fe(hipGraphExecDestroy(g->exec));
    if (g->graph) fe(hipGraphDestroy(g->graph));
    free_render(g->R);
    if (g->scratch) fe(hipFree(g->scratch));
    delete g;
    return fe.status("mt_render_graph_destroy");
}

extern "C" int32_t mt_render_graph_create(mt_context* ctx, const mt_scene* scene, const float* materials,
                                          const mt_envmap* env, const int32_t* light_faces,
                                          const int32_t* d_num_lights, const mt_render_desc* desc, float* image,
                                          mt_render_graph** out)
{
    MT_ARG_CHECK(ctx && out, "mt_render_graph_create: NULL argument");
    *out = nullptr;
    MT_ARG_CHECK(!desc || (desc->flags & MT_RENDER_PROFILE) == 0,
                 "mt_render_graph_create: MT_RENDER_PROFILE (event timing) cannot be captured");
    mt_render_graph* g = new (std::nothrow) mt_render_graph();
    if (!g) {
        mt_set_error("mt_render_graph_create: host allocation failed");
        return MT_ERR_ALLOC;
    }
    g->ctx = ctx;
    g->R.base = nullptr;
    g->R.nev = 0;
    g->R.fork = g->R.join = nullptr;
    int32_t rc = plan_render(scene, materials, env, light_faces, d_num_lights, desc, image, g->R);
    if (rc == MT_OK) rc = alloc_render(g->R, ctx, false);
    if (rc) {
        delete g;
        return rc;
    }
    // The graph gets its own look-back workspace: the context's may be
    // regrown (and freed) by later calls, while the graph keeps addresses.
    void* ctx_scratch = ctx->scratch;
    const size_t ctx_scratch_bytes = ctx->scratch_bytes;
    const hipStream_t ctx_stream = ctx->stream;
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    rc = mt_scan_reserve(ctx, g->R.np);
    g->scratch = ctx->scratch;
    g->scratch_bytes = ctx->scratch_bytes;
    // Capture on a stream of our own (the context's may be the null stream,
    // which cannot be captured); the workspace fill above is ordered before
    // every later launch of the graph on the context's stream.
    hipStream_t cs = nullptr;
    if (rc == MT_OK && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        mt_set_error("mt_render_graph_create: hipStreamCreate failed");
        rc = MT_ERR_HIP;
    }
    if (rc == MT_OK) {
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            mt_set_error("mt_render_graph_create: hipStreamBeginCapture failed: %s", hipGetErrorString(e));
            rc = MT_ERR_HIP;
        } else {
            ctx->stream = cs;
            rc = enqueue_render(ctx, g->R);
            ctx->stream = ctx_stream;
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(cs, &graph);
            g->graph = graph;
            if (rc == MT_OK && e != hipSuccess) {
                mt_set_error("mt_render_graph_create: hipStreamEndCapture failed: %s", hipGetErrorString(e));
                rc = MT_ERR_HIP;
            }
        }
    }
    ctx->stream = ctx_stream;
    ctx->scratch = ctx_scratch;
    ctx->scratch_bytes = ctx_scratch_bytes;
    if (cs) (void)hipStreamDestroy(cs);
    if (rc == MT_OK) {
        hipError_t e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            mt_set_error("mt_render_graph_create: hipGraphInstantiate failed: %s", hipGetErrorString(e));
            g->exec = nullptr;
            rc = MT_ERR_HIP;
        }
    }
    if (rc) {
        mt_render_graph_destroy(g);
        return rc;
    }
    *out = g;
    return MT_OK;
}

extern "C" int32_t mt_render_graph_launch(mt_render_graph* g, uint32_t seed)
{
    MT_ARG_CHECK(g && g->exec, "mt_render_graph_launch: NULL graph");
    hipStream_t st = g->ctx->stream;
    k_set_seed<<<1, 64, 0, st>>>(seed_ptr(g->R), seed);
    MT_LAUNCH_CHECK();
    MT_HIP_CHECK(hipGraphLaunch(g->exec, st));
    return MT_OK;
}

extern "C" int32_t mt_render_graph_stats(mt_render_graph* g, mt_render_stats* stats)
{
    MT_ARG_CHECK(g && stats, "mt_render_graph_stats: NULL argument");
    int32_t rc = read_stats(g->ctx, g->R, stats);
    if (rc) return rc;
    // the look-back error flag of the graph's own workspace (word 1)
    uint32_t flag = 0;
    uint32_t* tk = reinterpret_cast<uint32_t*>(g->scratch);
    MT_HIP_CHECK(hipMemcpyAsync(&flag, tk + 1, sizeof(flag), hipMemcpyDeviceToHost, g->ctx->stream));
    MT_HIP_CHECK(hipStreamSynchronize(g->ctx->stream));
    if (flag) {
        rc = mt_buf_fill_i32(g->ctx, reinterpret_cast<int32_t*>(tk + 1), 0, 1);
        if (rc) return rc;
        mt_set_error("render graph: scan/compaction look-back exceeded its spin limit; results invalid");
        return MT_ERR_STATE;
    }
    return MT_OK;
}
