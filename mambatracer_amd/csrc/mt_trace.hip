// Stage A: stackless BVH closest-hit / any-hit traversal on gfx950.
//
// Restates intersectMesh (src/cl/kernel/NormalBVH.cl:14-92), testOcclusion
// (:235-317), fastTriangleIntersection (Primitive.cl:3-29) and
// intersectBoundT (Util.cl:707-732) with the same visiting order: at each
// interior node both child boxes are tested, the nearer (entry distance,
// ties to the left) is taken first and the other is remembered by a 1 bit
// on a bitstack; backtracking walks parent links and flips to the sibling.
// Keeping that order (not just the result) keeps tie-breaks and the
// tMax-dependent culling identical, so primitive ids match bit for bit.
// The bitstack is 128 bits (two 64-bit words) instead of the reference's
// 64 (NormalBVH.cl:20), so trees up to depth 128 traverse correctly; scenes
// deeper than that are rejected at mt_scene_create.
//
// Two memory layouts:
//  * MT_TRACE_REFERENCE reads the reference's flat arrays (7-int nodes,
//    8-float boxes, 10-int faces, float4 points) as IntersectPrimitives does.
//  * MT_TRACE_PACKED (default) reads a layout derived once per scene: each
//    interior node is one 64-byte record holding both child boxes, both child
//    references and its own parent/sibling; each leaf triangle is one 48-byte
//    record (p1, e1 = p2 - p1, e2 = p3 - p1, primitive id, and the leaf's
//    parent/sibling in its two spare words) stored in leaf order.  e1/e2 are the same float subtractions the reference performs
//    per test, so every intersection rounds identically.  An interior step
//    fetches 64 B instead of 148 B, a leaf 48 B instead of 116 B.
#include "mt_traverse.h"

#include <cstdio>

namespace {

struct TraceArgs {
    const int* nodes;
    const float* bounds;
    const int* faces;
    const float4* points;
    const float4* pn;
    const float4* pt;
    int root;
    const float4* o_tmin;
    const float4* d_tmax;
    const int* d_count;
    int capacity;
    int* prim;
    float* tout;
    int* occluded;
    int* stats;
    unsigned* ctr; // refill kernel: [0] next ray, [1] waves finished; zero between launches
    MtSegQueue seg; // segmented queue (seg.counts set) instead of *d_count
    const int* order; // binned traversal: entry k is ray order[k], its results go there (null: ray k)
};

// Whether queue entry `id` holds a ray (plain queue: id < *d_count clamped
// to the capacity; segmented: below its segment's count).
__device__ __forceinline__ bool entry_live(const TraceArgs& A, int id)
{
    if (A.seg.counts) {
        const int p = id / A.seg.seg_cap;
        return p < A.seg.nseg && id - p * A.seg.seg_cap < A.seg.counts[p];
    }
    int count = *A.d_count;
    if (count > A.capacity) count = A.capacity;
    return id < count;
}

// Queue entry of this lane.  A segmented queue is walked chunk-major:
// work-group b takes chunk b / nseg of segment b % nseg, so the live chunks
// at the front of every segment are dispatched first, and neighbouring
// groups hold entries that neighbouring groups of the producer appended.
__device__ __forceinline__ int queue_entry(const TraceArgs& A)
{
    if (A.seg.counts) {
        const int p = blockIdx.x % A.seg.nseg, c = blockIdx.x / A.seg.nseg;
        return p * A.seg.seg_cap + c * blockDim.x + threadIdx.x;
    }
    return blockIdx.x * blockDim.x + threadIdx.x;
}

template <bool ANY, bool STATS, class State>
__device__ __forceinline__ void store_result(const TraceArgs& A, int id, const RayT& r, const State& S)
{
    if (A.order) id = A.order[id];
    if (ANY) {
        A.occluded[id] = S.hit;
    } else {
        A.prim[id] = S.hit ? S.child : -1;
        A.tout[id] = r.tmax;
    }
    if (STATS) {
        A.stats[3 * id + 0] = S.st[0];
        A.stats[3 * id + 1] = S.st[1];
        A.stats[3 * id + 2] = S.st[2];
    }
}

// Persistent traversal with per-lane ray refill.  Incoherent rays differ a
// lot in length, and a wave that owns 64 rays runs until its longest one
// ends.  Here a wave takes chunks of REFILL_CHUNK consecutive rays with one
// global atomic and, whenever at least REFILL_MIN of its lanes have finished
// (or all have), hands the chunk's next rays to the idle lanes (ballot +
// mbcnt), so lanes stay busy until the rays run out.  Each ray is traversed
// by trace_packed_step in the reference's visiting order; only the mapping of
// rays to lanes changes.  The last wave to leave clears the counters.
constexpr int REFILL_CHUNK = 1024;
constexpr int REFILL_MIN = 16;

template <bool ANY, bool STATS>
__global__ __launch_bounds__(256) void k_trace_refill(TraceArgs A)
{
    int count = *A.d_count;
    if (count > A.capacity) count = A.capacity;
    const int lane = __lane_id();
    int cnext = 0, cend = 0; // wave-uniform chunk cursor
    bool drained = false;    // wave-uniform: the global counter is past count
    bool active = false;
    int id = 0;
    RayT r;
    TravState S;
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (idle && !drained && (idle == ~0ull || __popcll(idle) >= REFILL_MIN)) {
            uint64_t want = idle;
            while (want) {
                if (cnext >= cend) {
                    unsigned base = 0;
                    if (lane == 0) base = atomicAdd(&A.ctr[0], (unsigned)REFILL_CHUNK);
                    base = __shfl(base, 0);
                    if (base >= (unsigned)count) {
                        drained = true;
                        break;
                    }
                    cnext = (int)base;
                    cend = (int)base + REFILL_CHUNK < count ? (int)base + REFILL_CHUNK : count;
                }
                int take = cend - cnext;
                int n = __popcll(want);
                take = take < n ? take : n;
                bool got = false;
                if ((want >> lane) & 1ull) {
                    int rk = (int)lane_rank(want);
                    if (rk < take) {
                        id = cnext + rk;
                        const int src = A.order ? A.order[id] : id;
                        ray_init(r, A.o_tmin[src], A.d_tmax[src]);
                        S.init(A.root);
                        active = true;
                        got = true;
                    }
                }
                cnext += take;
                want &= ~__ballot(got);
            }
        }
        if (!__any(active)) {
            if (drained) break;
            continue;
        }
        if (active && trace_packed_step<ANY, STATS, true>(A.pn, A.pt, r, S)) {
            store_result<ANY, STATS>(A, id, r, S);
            active = false;
        }
    }
    if (lane == 0) {
        unsigned waves = gridDim.x * (blockDim.x >> 6);
        unsigned done = atomicAdd(&A.ctr[1], 1u);
        if (done == waves - 1) {
            atomicExch(&A.ctr[0], 0u);
            atomicExch(&A.ctr[1], 0u);
        }
    }
}

// One ray per lane over a capacity-sized grid; slots >= *d_count exit at
// once (the hardware dispatcher retires such work-groups in microseconds,
// cheaper than a persistent grid drawing rays from one global counter).
// LAYOUT: 0 packed + short stack, 1 reference arrays, 2 packed without the
// short stack (the reference's parent-link pops, kept for comparison).
template <bool ANY, bool STATS, int LAYOUT>
__global__ __launch_bounds__(256) void k_trace(TraceArgs A)
{
    const int id = queue_entry(A);
    if (!entry_live(A, id)) return;
    RayT r;
    ray_init(r, A.o_tmin[id], A.d_tmax[id]);
    int st[3] = {0, 0, 0};
    int child = -1;
    int h = LAYOUT == 0   ? trace_packed<ANY, STATS, true>(A.pn, A.pt, A.root, r, child, st)
            : LAYOUT == 2 ? trace_packed<ANY, STATS, false>(A.pn, A.pt, A.root, r, child, st)
                          : trace_reference<ANY, STATS>(A.nodes, A.bounds, A.faces, A.points, r, child, st);
    if (ANY) {
        A.occluded[id] = h;
    } else {
        A.prim[id] = h ? child : -1;
        A.tout[id] = r.tmax;
    }
    if (STATS) {
        A.stats[3 * id + 0] = st[0];
        A.stats[3 * id + 1] = st[1];
        A.stats[3 * id + 2] = st[2];
    }
}

// Wave-pool traversal (pool_trace): each wave owns a run of 64 * POOL
// consecutive queue entries (segmented queue: work-group b takes chunk
// b / nseg of segment b % nseg, 256 * POOL entries, one run per wave; over a
// binned order, work-groups in per-XCD runs) and refills its finished lanes
// from it.
template <bool ANY, int POOL, int SS = MT_SHORT_STACK>
__global__ __launch_bounds__(256) void k_trace_pool(TraceArgs A)
{
    constexpr int RUN = 64 * POOL;
    int next, end;
    if (A.seg.counts) {
        pool_seg_run(A.seg.counts, A.seg.nseg, A.seg.seg_cap, RUN, next, end);
    } else {
        int count = *A.d_count;
        count = count < A.capacity ? count : A.capacity;
        count = count > 0 ? count : 0;
        // over a binned order, consecutive entries start in the same cell and
        // walk the same subtrees: the groups that hold entries are numbered
        // in per-XCD runs (the grid is sized for the capacity; the rest leave
        // at once below)
        const unsigned used = (unsigned)(((int64_t)count + 4 * RUN - 1) / (4 * RUN));
        const unsigned g = A.order && blockIdx.x < used ? xcd_chunked_group(blockIdx.x, used, XCD_RUN) : blockIdx.x;
        next = (int)(g * 4 + (threadIdx.x >> 6)) * RUN;
        end = count < next + RUN ? count : next + RUN;
    }
    if (next >= end) return;
    pool_trace<ANY, SS>(
        A.pn, A.pt, A.root, next, end,
        [&](int id, RayT& r, const float4*&, const float4*&, int&) {
            const int src = A.order ? A.order[id] : id; // binned: entry id is ray order[id]
            ray_init(r, A.o_tmin[src], A.d_tmax[src]);
        },
        [&](bool fin, int id, const RayT& r, const TravStateN<SS>& S) {
            if (fin) store_result<ANY, false>(A, id, r, S);
        });
}

// One ray per lane as k_trace LAYOUT 0, after the work-group's rays are
// ordered by direction: a ray's bin is its direction octant and dominant
// axis (24 bins, dead entries last), and lanes take the group's rays in bin
// order (stable within a bin).  A wave then holds rays heading the same way
// from the producer's neighbouring paths, which visit similar nodes, so
// fewer lanes idle while the wave's longest ray finishes.  Each ray is
// traversed as in k_trace and its result stored at its own entry: the
// output is identical, only the lane that computes it changes.
constexpr int SORT_BINS = 32;

__device__ __forceinline__ int dir_bin(float4 d)
{
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const int major = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const int oct = (d.x < 0.0f ? 1 : 0) | (d.y < 0.0f ? 2 : 0) | (d.z < 0.0f ? 4 : 0);
    return oct * 3 + major;
}

template <bool ANY>
__global__ __launch_bounds__(256) void k_trace_sorted(TraceArgs A)
{
    __shared__ int s_cnt[4][SORT_BINS];
    __shared__ int s_off[4][SORT_BINS];
    __shared__ int s_ids[256];
    const int id = queue_entry(A);
    const bool live = entry_live(A, id);
    if (__syncthreads_or(live) == 0) return; // no ray in this group
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    const int bin = live ? dir_bin(A.d_tmax[id]) : SORT_BINS - 1;
    int rank = 0;
#pragma unroll
    for (int b = 0; b < SORT_BINS; b++) {
        const uint64_t m = __ballot(bin == b);
        if (bin == b) rank = (int)lane_rank(m);
        if (lane == 0) s_cnt[wave][b] = (int)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < SORT_BINS) { // one lane per bin: exclusive scan of the bin totals
        const int b = threadIdx.x;
        const int c0 = s_cnt[0][b], c1 = s_cnt[1][b], c2 = s_cnt[2][b], c3 = s_cnt[3][b];
        const int tot = c0 + c1 + c2 + c3;
        int incl = tot;
        for (int o = 1; o < SORT_BINS; o <<= 1) {
            const int y = __shfl_up(incl, o, SORT_BINS);
            if (b >= o) incl += y;
        }
        const int ex = incl - tot;
        s_off[0][b] = ex;
        s_off[1][b] = ex + c0;
        s_off[2][b] = ex + c0 + c1;
        s_off[3][b] = ex + c0 + c1 + c2;
    }
    __syncthreads();
    s_ids[s_off[wave][bin] + rank] = live ? id : -1;
    __syncthreads();
    const int sid = s_ids[threadIdx.x];
    if (sid < 0) return;
    RayT r;
    ray_init(r, A.o_tmin[sid], A.d_tmax[sid]);
    int st[3] = {0, 0, 0};
    int child = -1;
    const int h = trace_packed<ANY, false, true>(A.pn, A.pt, A.root, r, child, st);
    if (ANY) {
        A.occluded[sid] = h;
    } else {
        A.prim[sid] = h ? child : -1;
        A.tout[sid] = r.tmax;
    }
}

// Rays of many jobs in one queue: ray id's scene is refs[tag[2*id] >> shift]
// (the job word of a wavefront slot).  Same per-ray traversal as k_trace
// LAYOUT 0, so a ray's result does not depend on which queue carries it.
__global__ __launch_bounds__(256) void k_trace_jobs(const MtSceneRef* __restrict__ refs, const int* __restrict__ tag,
                                                    int shift, TraceArgs A, const int* __restrict__ first)
{
    // first != nullptr: the entries from *first on (grid sized for all)
    const int id = (first ? *first : 0) + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (!entry_live(A, id)) return;
    const MtSceneRef R = refs[tag[2 * id] >> shift];
    RayT r;
    ray_init(r, A.o_tmin[id], A.d_tmax[id]);
    int st[3];
    int child = -1;
    int h = trace_packed<false, false, true>(R.pn, R.pt, R.root, r, child, st);
    A.prim[id] = h ? child : -1;
    A.tout[id] = r.tmax;
}

// ---------------------------------------------------------------- binned traversal
// Incoherent batches (random origins and directions) make neighbouring lanes
// and waves fetch unrelated nodes, and a BVH larger than the last-level
// cache then costs an HBM line per node visit.  MT_TRACE_PACKED_BINNED first
// orders the batch by the origin's cell in a 32 x 32 x 32 grid over the
// scene's root box (Morton order), with a one-pass counting sort (per-tile
// histograms in LDS, an exclusive scan of the bin-major tile x bin matrix,
// a scatter of the ray indices in key order), then traces the rays in that
// order with wave pools (k_trace_pool), each result stored at the ray's own index.
// Rays that start near each other visit much the same upper nodes while they
// are in flight together.  Per-ray results are those of any other mode.
// (64M random rays through 4M triangles, closest / any hit: 8^3 cells x
// direction octant 66.0 / 41.1 ms, 16^3 cells 63.8 / 39.0, 32^3 cells
// 63.5 / 38.6; 4^3 cells x 64 direction bins 67.3 / 41.3.)
constexpr int BIN_CELL_BITS = 5;                // cells per axis: 2^BIN_CELL_BITS
constexpr int BIN_CELLS = 1 << BIN_CELL_BITS;
constexpr int BIN_COUNT = 1 << (3 * BIN_CELL_BITS); // keys fit the 16-bit key array; LDS: 128 KB per group
// Rays per work-group (tile) of the histogram and scatter passes: the
// batch split into about one tile per CU (the histogram's LDS allows one
// work-group per CU), at least BIN_TILE_MIN rays, a multiple of 1024.
// Fewer, larger tiles shrink the tile x bin matrix (scattered 4-byte
// writes, its scan, scattered reads) and lengthen each (tile, bin) run of
// the order (64M rays, closest / any hit: 48.4 / 34.0 ms with 1024 tiles of
// 65,536, 48.0 / 33.4 ms with 256 of 262,144).
constexpr int BIN_TILE_MIN = 65536;
constexpr int BIN_TILES_TARGET = 256;
// threads per work-group of those passes: the 128 KB histogram allows one
// work-group per CU, so it brings its own waves to hide the LDS-atomic and
// store latency (64M rays: sort 2.7 ms per stage with 256 threads, step
// 86.8 -> 85.5 ms with 1024)
constexpr int BIN_THREADS = 1024;
constexpr int BINNED_POOL_RAYS = 8; // rays per lane of a wave's run when tracing the sorted order

__device__ __forceinline__ int bin_axis(float o, float lo, float scale)
{
    const float f = (o - lo) * scale; // NaN: cell 0
    return f >= 0.0f ? (f < (float)BIN_CELLS ? (int)f : BIN_CELLS - 1) : 0;
}

__device__ __forceinline__ int bin_key(float4 o, float4 d, const float* __restrict__ box, float sx, float sy,
                                       float sz)
{
    const int x = bin_axis(o.x, box[0], sx), y = bin_axis(o.y, box[1], sy), z = bin_axis(o.z, box[2], sz);
    int m = 0;
#pragma unroll
    for (int b = 0; b < BIN_CELL_BITS; b++)
        m |= (((x >> b) & 1) << (3 * b)) | (((y >> b) & 1) << (3 * b + 1)) | (((z >> b) & 1) << (3 * b + 2));
    (void)d;
    return m;
}

// The root box's per-axis scale cells / extent (0 for a flat axis).
__device__ __forceinline__ void bin_scales(const float* __restrict__ box, float& sx, float& sy, float& sz)
{
    const float ex = box[4] - box[0], ey = box[5] - box[1], ez = box[6] - box[2];
    sx = ex > 0.0f ? (float)BIN_CELLS / ex : 0.0f;
    sy = ey > 0.0f ? (float)BIN_CELLS / ey : 0.0f;
    sz = ez > 0.0f ? (float)BIN_CELLS / ez : 0.0f;
}

// Pass 1: key of every live ray, and each tile's key histogram written
// bin-major (hist[bin * ntiles + tile]) so one exclusive scan gives every
// (tile, bin) its first output position.
__global__ __launch_bounds__(BIN_THREADS) void k_bin_hist(const float4* __restrict__ o_tmin, const float4* __restrict__ d_tmax,
                                                  const int* __restrict__ d_count, int capacity,
                                                  const float* __restrict__ box, uint16_t* __restrict__ keys,
                                                  int* __restrict__ hist, int ntiles, int tile)
{
    __shared__ int h[BIN_COUNT];
    for (int i = threadIdx.x; i < BIN_COUNT; i += blockDim.x) h[i] = 0;
    __syncthreads();
    int count = *d_count;
    count = count < capacity ? count : capacity;
    float sx, sy, sz;
    bin_scales(box, sx, sy, sz);
    const int64_t base = (int64_t)blockIdx.x * tile;
    const int last = (int)(base + tile < (int64_t)count ? base + tile : (int64_t)count);
    for (int i = (int)base + (int)threadIdx.x; i < last; i += blockDim.x) {
        const int k = bin_key(o_tmin[i], d_tmax[i], box, sx, sy, sz);
        keys[i] = (uint16_t)k;
        atomicAdd(&h[k], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < BIN_COUNT; b += blockDim.x) hist[(size_t)b * ntiles + blockIdx.x] = h[b];
}

// Pass 2 (after the scan): each ray's index goes to its (tile, bin) range;
// rank inside the range by LDS atomics (the order within a range is free:
// every result is stored by original index).  The rays themselves stay where
// they are: the traversal reads each through the order, two 16-B loads
// against the kilobytes of nodes it then fetches.
__global__ __launch_bounds__(BIN_THREADS) void k_bin_scatter(const int* __restrict__ d_count, int capacity,
                                                     const uint16_t* __restrict__ keys, const int* __restrict__ start,
                                                     int ntiles, int tile, int* __restrict__ order)
{
    __shared__ int off[BIN_COUNT];
    for (int b = threadIdx.x; b < BIN_COUNT; b += blockDim.x) off[b] = start[(size_t)b * ntiles + blockIdx.x];
    __syncthreads();
    int count = *d_count;
    count = count < capacity ? count : capacity;
    const int64_t base = (int64_t)blockIdx.x * tile;
    const int last = (int)(base + tile < (int64_t)count ? base + tile : (int64_t)count);
    for (int i = (int)base + (int)threadIdx.x; i < last; i += blockDim.x) {
        order[atomicAdd(&off[keys[i]], 1)] = i;
    }
}

__global__ void k_set_i32(int* p, int v) { *p = v; }

// The multi-job queue with wave pools: each wave takes a run of 64 *
// POOL_RAYS consecutive slots and refills its lanes (pool_trace); a lane
// loads its slot's scene with its ray.
__global__ __launch_bounds__(256) void k_trace_jobs_pool(const MtSceneRef* __restrict__ refs,
                                                         const int* __restrict__ tag, int shift, TraceArgs A)
{
    int count = *A.d_count;
    count = count < A.capacity ? count : A.capacity;
    const int next = (int)(blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 * POOL_RAYS);
    const int end = count < next + 64 * POOL_RAYS ? count : next + 64 * POOL_RAYS;
    if (next >= end) return;
    pool_trace<false>(
        nullptr, nullptr, 0, next, end,
        [&](int id, RayT& r, const float4*& pn, const float4*& pt, int& root) {
            const MtSceneRef R = refs[tag[2 * id] >> shift];
            pn = R.pn;
            pt = R.pt;
            root = R.root;
            ray_init(r, A.o_tmin[id], A.d_tmax[id]);
        },
        [&](bool fin, int id, const RayT& r, const TravState& S) {
            if (fin) store_result<false, false>(A, id, r, S);
        });
}

template <bool ANY>
int32_t launch_trace(mt_context* ctx, const mt_scene* sc, const float* o_tmin,
                     const float* d_tmax, const int32_t* d_count, int32_t capacity,
                     int32_t* prim, float* t, int32_t* occluded, int32_t* stats, int32_t mode,
                     MtSegQueue seg = MtSegQueue{nullptr, 0, 0})
{
    MT_ARG_CHECK(ctx && sc && o_tmin && d_tmax && (d_count || seg.counts), "trace: NULL argument");
    MT_ARG_CHECK(capacity >= 0, "trace: negative capacity");
    MT_ARG_CHECK(mode >= MT_TRACE_PACKED && mode <= MT_TRACE_PACKED_BINNED, "trace: unknown mode %d", mode);
    if (ANY) MT_ARG_CHECK(occluded != nullptr, "trace_any: occluded is NULL");
    else MT_ARG_CHECK(prim && t, "trace_closest: prim/t NULL");
    if (seg.counts) {
        MT_ARG_CHECK(seg.seg_cap % 256 == 0, "trace: segment capacity %d not a multiple of 256", seg.seg_cap);
        capacity = seg.nseg * seg.seg_cap;
        if (mode == MT_TRACE_PACKED_REFILL) mode = MT_TRACE_PACKED_STACK; // the refill kernel reads *d_count
    }
    if (capacity == 0) return MT_OK;
    TraceArgs A{};
    A.seg = seg;
    A.nodes = sc->desc.nodes;
    A.bounds = sc->desc.bounds;
    A.faces = sc->desc.faces;
    A.points = reinterpret_cast<const float4*>(sc->desc.points);
    A.pn = sc->pnodes;
    A.pt = sc->ptris;
    A.root = sc->root_ref;
    A.o_tmin = reinterpret_cast<const float4*>(o_tmin);
    A.d_tmax = reinterpret_cast<const float4*>(d_tmax);
    A.d_count = d_count;
    A.capacity = capacity;
    A.prim = prim;
    A.tout = t;
    A.occluded = occluded;
    A.stats = stats;
    dim3 grid(mt_blocks(capacity, 256)), block(256);
    hipStream_t s = ctx->stream;
    // Default: one ray per lane with the short stack (8 waves/SIMD).  Refill
    // wins only on long incoherent batches (16M random any-hit rays through a
    // 4M-triangle soup: 18.3 vs 21.3 ms); on a path tracer's bounce batches it
    // loses (1080p x 16 spp render: 25.8 vs 23.5 ms), so it is opt-in.
    // On the render's segmented bounce queues (secondary rays, incoherent)
    // the wave pool wins (1080p x 16 spp guided render: bounce-1 closest
    // hit 2.80 -> 2.21 ms); on coherent batches its desynchronised lanes
    // fetch more distinct nodes per load and it loses (camera rays 1.40 ->
    // 1.90 ms), so plain queues keep one ray per lane.
    if (mode == MT_TRACE_PACKED) mode = seg.counts ? MT_TRACE_PACKED_POOL : MT_TRACE_PACKED_STACK;
    if ((mode == MT_TRACE_PACKED_SORTED || mode == MT_TRACE_PACKED_POOL || mode == MT_TRACE_PACKED_BINNED) && stats)
        mode = MT_TRACE_PACKED_STACK; // counting: per-ray order only
    if (mode == MT_TRACE_PACKED_BINNED && seg.counts) mode = MT_TRACE_PACKED_POOL; // the render's bounce queues
    if (mode == MT_TRACE_PACKED_BINNED) {
        // workspace: keys, tile x bin matrix and its scan, the order
        int64_t tile = ((int64_t)capacity + BIN_TILES_TARGET - 1) / BIN_TILES_TARGET;
        tile = tile < BIN_TILE_MIN ? BIN_TILE_MIN : (tile + 1023) & ~(int64_t)1023;
        const int64_t ntiles = (capacity + tile - 1) / tile;
        const int64_t cells = ntiles * BIN_COUNT;
        MT_ARG_CHECK(cells < (1ll << 31), "trace: %d rays too many for the binned mode", capacity);
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t b_keys = up(2 * (size_t)capacity), b_hist = up(4 * (size_t)cells);
        const size_t b_idx = up(4 * (size_t)capacity);
        void* w = nullptr;
        int32_t rc = mt_ctx_temp(ctx, b_keys + 2 * b_hist + b_idx + 256, &w);
        if (rc) return rc;
        char* p = reinterpret_cast<char*>(w);
        uint16_t* keys = reinterpret_cast<uint16_t*>(p);
        int* hist = reinterpret_cast<int*>(p + b_keys);
        int* start = reinterpret_cast<int*>(p + b_keys + b_hist);
        int* order = reinterpret_cast<int*>(p + b_keys + 2 * b_hist);
        int* nd = reinterpret_cast<int*>(p + b_keys + 2 * b_hist + b_idx); // [0] cells, [1] total
        const float* box = sc->desc.bounds; // the root node's box (CNormalBVH: node 0 owns bound 0)
        k_set_i32<<<1, 1, 0, s>>>(nd, (int)cells);
        k_bin_hist<<<(unsigned)ntiles, BIN_THREADS, 0, s>>>(A.o_tmin, A.d_tmax, A.d_count, capacity, box, keys, hist,
                                                   (int)ntiles, (int)tile);
        MT_LAUNCH_CHECK();
        rc = mt_scan_excl_i32(ctx, hist, start, nd, (int)cells, nd + 1);
        if (rc) return rc;
        k_bin_scatter<<<(unsigned)ntiles, BIN_THREADS, 0, s>>>(A.d_count, capacity, keys, start, (int)ntiles,
                                                      (int)tile, order);
        MT_LAUNCH_CHECK();
        A.order = order;
        // wave pools over the sorted order (64M random rays through 4M
        // triangles, closest / any hit: 51.0 / 35.7 ms; persistent refill
        // waves over the same order 63.9 / 39.0 ms): each wave's run is 256
        // neighbouring rays, whose lanes the pool keeps on nodes
        mode = MT_TRACE_PACKED_POOL;
    }
    if (mode == MT_TRACE_PACKED_POOL) {
        // sorted (binned) batches take runs of 8 rays per lane (64M random
        // rays, closest / any hit: 49.9 / 34.8 ms against 51.0 / 35.8 with
        // 4) and a deeper short stack; the render's bounce queues keep 4
        // rays and 8 entries (8 rays cost its job 1%, 12 entries 0.7%)
        const int rays = A.order ? BINNED_POOL_RAYS : POOL_RAYS;
        const int64_t per_group = 256 * (int64_t)rays;
        const int64_t groups = seg.counts ? (int64_t)seg.nseg * ((seg.seg_cap + per_group - 1) / per_group)
                                          : (capacity + per_group - 1) / per_group;
        if (A.order)
            k_trace_pool<ANY, BINNED_POOL_RAYS, SHORT_STACK_DEEP><<<dim3((unsigned)groups), block, 0, s>>>(A);
        else k_trace_pool<ANY, POOL_RAYS><<<dim3((unsigned)groups), block, 0, s>>>(A);
    } else if (mode == MT_TRACE_PACKED_SORTED) {
        k_trace_sorted<ANY><<<grid, block, 0, s>>>(A);
    } else if (mode == MT_TRACE_PACKED_REFILL) {
        void* scr = nullptr;
        int32_t rc = mt_ctx_scratch(ctx, 256, &scr);
        if (rc) return rc;
        // words 0-2: scan look-back ticket, error flag and epoch
        A.ctr = reinterpret_cast<unsigned*>(scr) + 4;
        auto kern = stats ? k_trace_refill<ANY, true> : k_trace_refill<ANY, false>;
        int blocks = mt_persistent_blocks(ctx, kern, 256, capacity);
        kern<<<blocks, block, 0, s>>>(A);
    } else if (stats) {
        if (mode == MT_TRACE_PACKED_STACK) k_trace<ANY, true, 0><<<grid, block, 0, s>>>(A);
        else if (mode == MT_TRACE_REFERENCE) k_trace<ANY, true, 1><<<grid, block, 0, s>>>(A);
        else k_trace<ANY, true, 2><<<grid, block, 0, s>>>(A);
    } else {
        if (mode == MT_TRACE_PACKED_STACK) k_trace<ANY, false, 0><<<grid, block, 0, s>>>(A);
        else if (mode == MT_TRACE_REFERENCE) k_trace<ANY, false, 1><<<grid, block, 0, s>>>(A);
        else k_trace<ANY, false, 2><<<grid, block, 0, s>>>(A);
    }
    MT_LAUNCH_CHECK();
    return MT_OK;
}

// -------------------------------------------------------------------------
// Scene preparation
// -------------------------------------------------------------------------
__global__ void k_leaf_flags(const int* __restrict__ nodes, int n, int* __restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = nodes[7 * i + 5] ? 1 : 0;
}

// Range checks of every index the traversal follows.  err bit 1: node link
// out of range; 2: face vertex out of range; 4: leaf primitive out of range;
// 8: bound index != own index (the reference's invariant, CNormalBVH.java:88).
__global__ void k_validate(const int* __restrict__ nodes, int nn, const int* __restrict__ faces,
                           int ntris, int npoints, int* err)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int e = 0;
    if (i < nn) {
        const int* n = nodes + 7 * i;
        if (n[0] != i) e |= 8;
        if (n[5]) {
            if (n[6] < 0 || n[6] >= ntris) e |= 4;
        } else {
            if (n[2] <= 0 || n[2] >= nn || n[3] <= 0 || n[3] >= nn) e |= 1;
        }
        if (n[1] < 0 || n[1] >= nn || n[4] < 0 || n[4] >= nn) e |= 1;
    }
    if (i < ntris) {
        const int* F = faces + 10 * i;
        for (int k = 0; k < 3; k++)
            if (F[k] < 0 || F[k] >= npoints) e |= 2;
    }
    if (e) atomicOr(err, e);
}

// Leaf depth by walking parent links (bounded): err |= 16 when a leaf is
// deeper than 128 or its parent chain does not reach the root.
__global__ void k_depth(const int* __restrict__ nodes, int nn, int* err)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn || !nodes[7 * i + 5] || i == 0) return;
    int d = 0, cur = i;
    while (cur != 0 && d <= 128) {
        cur = nodes[7 * cur + 4];
        d++;
    }
    if (cur != 0 || d > 128) atomicOr(err, 16);
}

__global__ void k_pack(const int* __restrict__ nodes, const float* __restrict__ bounds,
                       const int* __restrict__ faces, const float4* __restrict__ points, int nn,
                       const int* __restrict__ leaf_rank, float4* __restrict__ pn,
                       float4* __restrict__ pt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int* n = nodes + 7 * i;
    auto ref = [&](int j) -> int {
        return nodes[7 * j + 5] ? ~leaf_rank[j] : j - leaf_rank[j];
    };
    if (n[5]) {
        int c = n[6];
        const int* F = faces + 10 * c;
        float4 q1 = points[F[0]], q2 = points[F[1]], q3 = points[F[2]];
        v3 p1 = xyz(q1), e1 = sub3(xyz(q2), p1), e2 = sub3(xyz(q3), p1);
        float4* T = pt + 3 * (size_t)leaf_rank[i];
        T[0] = make_float4(p1.x, p1.y, p1.z, e1.x);
        T[1] = make_float4(e1.y, e1.z, e2.x, e2.y);
        // spare words: the leaf's parent record and sibling reference (node_links)
        const int parent = n[4];
        T[2] = make_float4(e2.z, __int_as_float(c), __int_as_float(parent - leaf_rank[parent]),
                           __int_as_float(ref(n[1])));
    } else {
        const float* lb = bounds + 8 * nodes[7 * n[2]];
        const float* rb = bounds + 8 * nodes[7 * n[3]];
        float4* N = pn + 4 * (size_t)(i - leaf_rank[i]);
        N[0] = make_float4(lb[0], lb[1], lb[2], lb[4]);
        N[1] = make_float4(lb[5], lb[6], rb[0], rb[1]);
        N[2] = make_float4(rb[2], rb[4], rb[5], rb[6]);
        int parent = n[4];
        int4 d = make_int4(ref(n[2]), ref(n[3]), parent - leaf_rank[parent], ref(n[1]));
        N[3] = *reinterpret_cast<float4*>(&d);
    }
}

// -------------------------------------------------------------------------
// Hit shading (IntersectPrimitives, NormalBVH.cl:124-232)
// -------------------------------------------------------------------------
__device__ __forceinline__ void barycentrics(v3 p, v3 p1, v3 p2, v3 p3, float& b1, float& b2)
{ // triangleBarycentrics, Primitive.cl:31-53
    v3 e1 = sub3(p2, p1), e2 = sub3(p3, p1), e = sub3(p, p1);
    float d00 = dot3(e1, e1), d01 = dot3(e1, e2), d11 = dot3(e2, e2);
    float d20 = dot3(e, e1), d21 = dot3(e, e2);
    float denom = (d00 * d11 - d01 * d01);
    if (denom == 0.0f) {
        b1 = 0.0f;
        b2 = 0.0f;
        return;
    }
    float inv = 1.0f / denom;
    b1 = (d11 * d20 - d01 * d21) * inv;
    b2 = (d00 * d21 - d01 * d20) * inv;
}

// Hit record of IntersectPrimitives (NormalBVH.cl:124-220) for ray (o, d)
// whose closest hit is primitive c at distance tt.  Every w lane is 0: the
// reference derives n and ng from w = 0 inputs, and leaves p.w unset.
struct HitShading {
    float4 p, n, ng;
    float2 uv;
    int mat;
};

__device__ __forceinline__ HitShading hit_shading(const float4* __restrict__ points,
                                                  const float4* __restrict__ normals,
                                                  const float2* __restrict__ uvs,
                                                  const int* __restrict__ faces, v3 o, v3 d, int c,
                                                  float tt)
{
    const int* F = faces + 10 * c;
    v3 p1 = xyz(points[F[0]]), p2 = xyz(points[F[1]]), p3 = xyz(points[F[2]]);
    v3 p = mk3(o.x + tt * d.x, o.y + tt * d.y, o.z + tt * d.z); // getPoint, Util.cl:558-565
    float texu = 0.0f, texv = 0.0f;
    bool has_uv = F[3] > -1 && uvs;
    if (has_uv) { // triangleBarycentricsFromUVMesh, Primitive.cl:56-71
        v3 f1 = sub3(p1, p), f2 = sub3(p2, p), f3 = sub3(p3, p);
        float a = len3(cross3(sub3(p1, p2), sub3(p1, p3)));
        float a1 = len3(cross3(f2, f3)) / a;
        float a2 = len3(cross3(f3, f1)) / a;
        float a3 = len3(cross3(f1, f2)) / a;
        float2 u1 = uvs[F[3]], u2 = uvs[F[4]], u3 = uvs[F[5]];
        texu = (u1.x * a1 + u2.x * a2) + u3.x * a3;
        texv = (u1.y * a1 + u2.y * a2) + u3.y * a3;
    }
    float u, v;
    barycentrics(p, p1, p2, p3, u, v);
    v3 n;
    if (F[6] > -1 && normals) {
        v3 n1 = xyz(normals[F[6]]), n2 = xyz(normals[F[7]]), n3 = xyz(normals[F[8]]);
        float w0 = (1.0f - u) - v;
        n = add3(add3(scl3(n1, w0), scl3(n2, u)), scl3(n3, v));
        if (dot3(n, d) > 0.00001f) {
            n = neg3(n);
            n1 = neg3(n1);
            n2 = neg3(n2);
            n3 = neg3(n3);
        }
        n = norm3(n);
        // Hanika shadow-terminator offset, NormalBVH.cl:175-193
        v3 tu = sub3(p, p1), tv = sub3(p, p2), tw = sub3(p, p3);
        v3 nn1 = norm3(n1), nn2 = norm3(n2), nn3 = norm3(n3);
        float du = fminf(0.0f, dot3(tu, nn1));
        float dv = fminf(0.0f, dot3(tv, nn2));
        float dw = fminf(0.0f, dot3(tw, nn3));
        tu = sub3(tu, scl3(nn1, du));
        tv = sub3(tv, scl3(nn2, dv));
        tw = sub3(tw, scl3(nn3, dw));
        p = add3(add3(add3(p, scl3(tu, w0)), scl3(tv, u)), scl3(tw, v));
    } else {
        n = norm3(cross3(sub3(p2, p1), sub3(p3, p1))); // getNormal, Util.cl:349-357
        if (dot3(n, d) > 0.00001f) n = neg3(n);
    }
    v3 ng = norm3(cross3(sub3(p2, p1), sub3(p3, p1)));
    if (dot3(ng, d) > 0.00001f) ng = neg3(ng);
    HitShading H;
    H.p = make_float4(p.x, p.y, p.z, 0.0f);
    H.n = make_float4(n.x, n.y, n.z, 0.0f);
    H.ng = make_float4(ng.x, ng.y, ng.z, 0.0f);
    H.uv = has_uv ? make_float2(texu, texv) : make_float2(u, v);
    H.mat = F[9] & 0xFFFF;
    return H;
}

__global__ __launch_bounds__(256) void k_shade(const float4* __restrict__ points,
                                               const float4* __restrict__ normals,
                                               const float2* __restrict__ uvs,
                                               const int* __restrict__ faces,
                                               const float4* __restrict__ o_tmin,
                                               const float4* __restrict__ d_tmax,
                                               const int* __restrict__ prim,
                                               const float* __restrict__ tin,
                                               const int* __restrict__ d_count, int capacity,
                                               float4* p_out, float4* n_out, float4* ng_out,
                                               float2* uv_out, int* mat_out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int count = *d_count;
    if (count > capacity) count = capacity;
    if (i >= count) return;
    int c = prim[i];
    if (c < 0) {
        mat_out[i] = -1;
        float4 z = make_float4(0, 0, 0, 0);
        p_out[i] = z;
        n_out[i] = z;
        ng_out[i] = z;
        uv_out[i] = make_float2(0, 0);
        return;
    }
    float4 dd = d_tmax[i], oo = o_tmin[i];
    const HitShading H = hit_shading(points, normals, uvs, faces, mk3(oo.x, oo.y, oo.z), mk3(dd.x, dd.y, dd.z), c,
                                     tin[i]);
    p_out[i] = H.p;
    n_out[i] = H.n;
    ng_out[i] = H.ng;
    uv_out[i] = H.uv;
    mat_out[i] = H.mat;
}

// IntersectPrimitives over the reference's own records (NormalBVH.cl:94-233).
// A Ray is 80 bytes (Util.cl:129-138): o, d, inv_d (float4 each), sign
// (int4), extra (int2), tMin, tMax, so as float4s: o, d, inv_d, sign,
// (extra.x, extra.y, tMin, tMax).  An Intersection is 96 bytes
// (Util.cl:141-151): p, n, ng, d (float4 each), uv (float2), mat, id, hit,
// 12 bytes of padding.  The box tests use the record's inv_d and sign, as
// the reference does.  A hit writes the whole intersection and the hit
// distance into the ray's tMax (intersectMesh, NormalBVH.cl:42); a miss,
// and every slot in [*count, capacity), writes hit = 0 and mat = -1 and
// leaves the other fields as they were.
template <int LAYOUT>
__global__ __launch_bounds__(256) void k_intersect_records(TraceArgs A, const float4* __restrict__ normals,
                                                           const float2* __restrict__ uvs, float4* rays,
                                                           float4* isects)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= A.capacity) return;
    int* iw = reinterpret_cast<int*>(isects + 6 * (size_t)id);
    float4* R = rays + 5 * (size_t)id;
    int child = -1;
    int h = 0;
    RayT r;
    float4 d4 = make_float4(0, 0, 0, 0);
    if (entry_live(A, id)) {
        const float4 o4 = R[0], inv4 = R[2], sg = R[3], ex = R[4];
        d4 = R[1];
        r.o = mk3(o4.x, o4.y, o4.z);
        r.d = mk3(d4.x, d4.y, d4.z);
        r.inv = mk3(inv4.x, inv4.y, inv4.z);
        r.sx = __float_as_int(sg.x) != 0;
        r.sy = __float_as_int(sg.y) != 0;
        r.sz = __float_as_int(sg.z) != 0;
        r.tmin = ex.z;
        r.tmax = ex.w;
        int st[3];
        h = LAYOUT == 0   ? trace_packed<false, false, true>(A.pn, A.pt, A.root, r, child, st)
            : LAYOUT == 2 ? trace_packed<false, false, false>(A.pn, A.pt, A.root, r, child, st)
                          : trace_reference<false, false>(A.nodes, A.bounds, A.faces, A.points, r, child, st);
    }
    if (!h) {
        iw[18] = -1; // mat
        iw[20] = 0;  // hit = MISS_MARKER
        return;
    }
    reinterpret_cast<float*>(R)[19] = r.tmax;
    const HitShading H = hit_shading(A.points, normals, uvs, A.faces, r.o, r.d, child, r.tmax);
    float4* I = isects + 6 * (size_t)id;
    I[0] = H.p;
    I[1] = H.n;
    I[2] = H.ng;
    I[3] = d4;
    I[4] = make_float4(H.uv.x, H.uv.y, __int_as_float(H.mat), __int_as_float(child));
    iw[20] = 1; // hit = HIT_MARKER
}

// intersectOcclusion (NormalBVH.cl:319-348) over Ray records: intersectMesh
// with bailout, i.e. any hit in the reference's visiting order, the ray left
// unchanged; hits[id] written for id < *count only.
template <int LAYOUT>
__global__ __launch_bounds__(256) void k_occlusion_records(TraceArgs A, const float4* __restrict__ rays, int* hits)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (!entry_live(A, id)) return;
    const float4* R = rays + 5 * (size_t)id;
    const float4 o4 = R[0], d4 = R[1], inv4 = R[2], sg = R[3], ex = R[4];
    RayT r;
    r.o = mk3(o4.x, o4.y, o4.z);
    r.d = mk3(d4.x, d4.y, d4.z);
    r.inv = mk3(inv4.x, inv4.y, inv4.z);
    r.sx = __float_as_int(sg.x) != 0;
    r.sy = __float_as_int(sg.y) != 0;
    r.sz = __float_as_int(sg.z) != 0;
    r.tmin = ex.z;
    r.tmax = ex.w;
    int st[3];
    int child = -1;
    hits[id] = LAYOUT == 0   ? trace_packed<true, false, true>(A.pn, A.pt, A.root, r, child, st)
               : LAYOUT == 2 ? trace_packed<true, false, false>(A.pn, A.pt, A.root, r, child, st)
                             : trace_reference<true, false>(A.nodes, A.bounds, A.faces, A.points, r, child, st);
}

__global__ __launch_bounds__(256) void k_shadow_rays(const float4* __restrict__ o_tmin,
                                                     const float4* __restrict__ d_tmax,
                                                     const int* __restrict__ prim,
                                                     const float* __restrict__ tin, float lx,
                                                     float ly, float lz,
                                                     const int* __restrict__ d_count,
                                                     int capacity, float4* so, float4* sd,
                                                     int* valid)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int count = *d_count;
    if (count > capacity) count = capacity;
    if (i >= count) return;
    bool ok = prim[i] >= 0;
    float tt = ok ? tin[i] : 0.0f;
    float4 o = o_tmin[i], d = d_tmax[i];
    float px = o.x + tt * d.x, py = o.y + tt * d.y, pz = o.z + tt * d.z;
    float vx = lx - px, vy = ly - py, vz = lz - pz;
    float dist = sqrtf((vx * vx + vy * vy) + vz * vz);
    so[i] = make_float4(px, py, pz, 0.0001f);
    sd[i] = make_float4(vx / dist, vy / dist, vz / dist, dist - 2 * MT_EPS_RAY);
    valid[i] = ok;
}

} // namespace

extern "C" int32_t mt_scene_create(mt_context* ctx, const mt_scene_desc* desc, mt_scene** out)
{
    MT_ARG_CHECK(ctx && desc && out, "mt_scene_create: NULL argument");
    *out = nullptr;
    MT_ARG_CHECK(desc->num_tris >= 1 && desc->num_points >= 3, "mt_scene_create: empty scene");
    MT_ARG_CHECK(desc->num_tris <= (1 << 30), "mt_scene_create: too many triangles");
    MT_ARG_CHECK(desc->points && desc->faces && desc->nodes && desc->bounds,
                 "mt_scene_create: missing scene array");
    if (desc->max_depth > 128) {
        mt_set_error("mt_scene_create: BVH depth %d exceeds the 128-level traversal bitstack",
                     desc->max_depth);
        return MT_ERR_DEPTH;
    }
    int nn = 2 * desc->num_tris - 1;
    int ninterior = desc->num_tris - 1;
    mt_scene* sc = new mt_scene();
    sc->ctx = ctx;
    sc->desc = *desc;
    sc->num_nodes = nn;
    sc->num_interior = ninterior;
    sc->pnodes = nullptr;
    sc->ptris = nullptr;
    int* work = nullptr; // [nn] flags, [nn] ranks, [2] err/total
    auto fail = [&](int32_t code) {
        if (work) (void)hipFree(work);
        if (sc->pnodes) (void)hipFree(sc->pnodes);
        if (sc->ptris) (void)hipFree(sc->ptris);
        delete sc;
        return code;
    };
    if (hipMalloc(&work, sizeof(int) * (2 * (size_t)nn + 4)) != hipSuccess ||
        hipMalloc(&sc->pnodes, sizeof(float4) * 4 * (size_t)(ninterior > 0 ? ninterior : 1)) != hipSuccess ||
        hipMalloc(&sc->ptris, sizeof(float4) * 3 * (size_t)desc->num_tris) != hipSuccess) {
        mt_set_error("mt_scene_create: device allocation failed (%d triangles)", desc->num_tris);
        return fail(MT_ERR_ALLOC);
    }
    int* flags = work;
    int* ranks = work + nn;
    int* err = work + 2 * (size_t)nn;
    int* total = err + 1;
    hipStream_t s = ctx->stream;
    int32_t rc = mt_buf_fill_i32(ctx, err, 0, 2);
    if (rc) return fail(rc);
    k_validate<<<mt_blocks(nn, 256), 256, 0, s>>>(desc->nodes, nn, desc->faces, desc->num_tris,
                                                  desc->num_points, err);
    int herr = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        mt_set_error("mt_scene_create: validation launch failed");
        return fail(MT_ERR_HIP);
    }
    if (herr) {
        mt_set_error("mt_scene_create: invalid scene arrays (error mask 0x%x: 1 node link, "
                     "2 face vertex, 4 leaf primitive, 8 bound index)", herr);
        return fail(MT_ERR_ARG);
    }
    k_depth<<<mt_blocks(nn, 256), 256, 0, s>>>(desc->nodes, nn, err);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        mt_set_error("mt_scene_create: depth check launch failed");
        return fail(MT_ERR_HIP);
    }
    if (herr) {
        mt_set_error("mt_scene_create: BVH deeper than 128 levels or parent links broken");
        return fail(MT_ERR_DEPTH);
    }
    k_leaf_flags<<<mt_blocks(nn, 256), 256, 0, s>>>(desc->nodes, nn, flags);
    rc = mt_buf_fill_i32(ctx, total, nn, 1); // device-side length for the scan
    if (rc) return fail(rc);
    rc = mt_scan_excl_i32(ctx, flags, ranks, total, nn, total + 1);
    if (rc) return fail(rc);
    k_pack<<<mt_blocks(nn, 256), 256, 0, s>>>(desc->nodes, desc->bounds, desc->faces,
                                              reinterpret_cast<const float4*>(desc->points), nn,
                                              ranks, sc->pnodes, sc->ptris);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        mt_set_error("mt_scene_create: packing failed");
        return fail(MT_ERR_HIP);
    }
    (void)hipFree(work);
    sc->root_ref = desc->num_tris == 1 ? ~0 : 0;
    *out = sc;
    return MT_OK;
}

extern "C" int32_t mt_scene_destroy(mt_scene* sc)
{
    if (!sc) return MT_OK;
    MtFirstError fe;
    fe(hipStreamSynchronize(sc->ctx->stream));
    if (sc->pnodes) fe(hipFree(sc->pnodes));
    if (sc->ptris) fe(hipFree(sc->ptris));
    delete sc;
    return fe.status("mt_scene_destroy");
}

extern "C" int32_t mt_trace_closest(mt_context* ctx, const mt_scene* scene, const float* o_tmin,
                                    const float* d_tmax, const int32_t* d_count, int32_t capacity,
                                    int32_t* prim, float* t, int32_t* stats, int32_t mode)
{
    return launch_trace<false>(ctx, scene, o_tmin, d_tmax, d_count, capacity, prim, t, nullptr,
                               stats, mode);
}

extern "C" int32_t mt_trace_any(mt_context* ctx, const mt_scene* scene, const float* o_tmin,
                                const float* d_tmax, const int32_t* d_count, int32_t capacity,
                                int32_t* occluded, int32_t* stats, int32_t mode)
{
    return launch_trace<true>(ctx, scene, o_tmin, d_tmax, d_count, capacity, nullptr, nullptr,
                              occluded, stats, mode);
}

extern "C" int32_t mt_trace_autotune(mt_context* ctx, const mt_scene* scene, const float* o_tmin,
                                     const float* d_tmax, const int32_t* d_count, int32_t capacity, int32_t any,
                                     int32_t* prim, float* t, int32_t* occluded, int32_t reps, int32_t* best_mode,
                                     float* best_ms)
{
    MT_ARG_CHECK(ctx && scene && best_mode && best_ms, "mt_trace_autotune: NULL argument");
    MT_ARG_CHECK(reps >= 1 && reps <= 100, "mt_trace_autotune: reps %d not in [1,100]", reps);
    MT_ARG_CHECK(any ? occluded != nullptr : (prim && t), "mt_trace_autotune: NULL outputs");
    const int32_t modes[4] = {MT_TRACE_PACKED_STACK, MT_TRACE_PACKED_REFILL, MT_TRACE_PACKED_POOL,
                              MT_TRACE_PACKED_BINNED};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MT_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        mt_set_error("mt_trace_autotune: hipEventCreate failed");
        return MT_ERR_HIP;
    }
    auto run = [&](int32_t mode) -> int32_t {
        return any ? launch_trace<true>(ctx, scene, o_tmin, d_tmax, d_count, capacity, nullptr, nullptr, occluded,
                                        nullptr, mode)
                   : launch_trace<false>(ctx, scene, o_tmin, d_tmax, d_count, capacity, prim, t, nullptr, nullptr,
                                         mode);
    };
    int32_t rc = MT_OK;
    float best = -1.0f;
    int32_t pick = modes[0];
    for (int32_t mode : modes) {
        rc = run(mode); // warm-up
        if (rc) break;
        hipError_t e = hipEventRecord(e0, ctx->stream);
        for (int k = 0; k < reps && rc == MT_OK; k++) rc = run(mode);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            mt_set_error("mt_trace_autotune: timing failed: %s", hipGetErrorString(e));
            rc = MT_ERR_HIP;
            break;
        }
        ms /= (float)reps;
        if (best < 0.0f || ms < best) {
            best = ms;
            pick = mode;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *best_mode = pick;
    *best_ms = best;
    return MT_OK;
}

int32_t mt_trace_closest_seg(mt_context* ctx, const mt_scene* scene, const float4* o_tmin, const float4* d_tmax,
                             MtSegQueue seg, int32_t* prim, float* t, int32_t mode)
{
    MT_ARG_CHECK(seg.counts && seg.nseg > 0 && seg.seg_cap > 0, "trace_closest_seg: bad segment table");
    return launch_trace<false>(ctx, scene, reinterpret_cast<const float*>(o_tmin),
                               reinterpret_cast<const float*>(d_tmax), nullptr, 0, prim, t, nullptr, nullptr, mode,
                               seg);
}

int32_t mt_trace_any_seg(mt_context* ctx, const mt_scene* scene, const float4* o_tmin, const float4* d_tmax,
                         MtSegQueue seg, int32_t* occluded, int32_t* stats, int32_t mode)
{
    MT_ARG_CHECK(seg.counts && seg.nseg > 0 && seg.seg_cap > 0, "trace_any_seg: bad segment table");
    return launch_trace<true>(ctx, scene, reinterpret_cast<const float*>(o_tmin),
                              reinterpret_cast<const float*>(d_tmax), nullptr, 0, nullptr, nullptr, occluded, stats,
                              mode, seg);
}

int32_t mt_trace_jobs(mt_context* ctx, const MtSceneRef* refs, const int32_t* tag, int32_t shift,
                      const float4* o_tmin, const float4* d_tmax, const int32_t* d_count, int32_t capacity,
                      int32_t* prim, float* t, int32_t mode, const int32_t* d_split)
{
    if (capacity == 0) return MT_OK;
    TraceArgs A{};
    A.o_tmin = o_tmin;
    A.d_tmax = d_tmax;
    A.d_count = d_count;
    A.capacity = capacity;
    A.prim = prim;
    A.tout = t;
    if (mode == MT_BATCH_TRACE_SPLIT) {
        MT_ARG_CHECK(d_split, "mt_trace_jobs: the split mode needs the survivors' count");
        // survivors (incoherent secondary rays) with wave pools ...
        TraceArgs P = A;
        P.d_count = d_split;
        k_trace_jobs_pool<<<mt_blocks(capacity, 256 * POOL_RAYS), 256, 0, ctx->stream>>>(refs, tag, shift, P);
        MT_LAUNCH_CHECK();
        // ... and the new camera rays (coherent) one per lane
        k_trace_jobs<<<mt_blocks(capacity, 256), 256, 0, ctx->stream>>>(refs, tag, shift, A, d_split);
    } else if (mode == MT_TRACE_PACKED_POOL) {
        k_trace_jobs_pool<<<mt_blocks(capacity, 256 * POOL_RAYS), 256, 0, ctx->stream>>>(refs, tag, shift, A);
    } else {
        k_trace_jobs<<<mt_blocks(capacity, 256), 256, 0, ctx->stream>>>(refs, tag, shift, A, nullptr);
    }
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_shade_hits(mt_context* ctx, const mt_scene* sc, const float* o_tmin,
                                 const float* d_tmax, const int32_t* prim, const float* t,
                                 const int32_t* d_count, int32_t capacity, float* p, float* n,
                                 float* ng, float* uv, int32_t* mat)
{
    MT_ARG_CHECK(ctx && sc && o_tmin && d_tmax && prim && t && d_count && p && n && ng && uv && mat,
                 "mt_shade_hits: NULL argument");
    if (capacity <= 0) return MT_OK;
    k_shade<<<mt_blocks(capacity, 256), 256, 0, ctx->stream>>>(
        reinterpret_cast<const float4*>(sc->desc.points),
        reinterpret_cast<const float4*>(sc->desc.normals),
        reinterpret_cast<const float2*>(sc->desc.uvs), sc->desc.faces,
        reinterpret_cast<const float4*>(o_tmin), reinterpret_cast<const float4*>(d_tmax), prim, t,
        d_count, capacity, reinterpret_cast<float4*>(p), reinterpret_cast<float4*>(n),
        reinterpret_cast<float4*>(ng), reinterpret_cast<float2*>(uv), mat);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_intersect_primitives(mt_context* ctx, const mt_scene* sc, void* rays, void* isects,
                                           const int32_t* d_count, int32_t capacity, int32_t mode)
{
    MT_ARG_CHECK(ctx && sc && rays && isects && d_count, "mt_intersect_primitives: NULL argument");
    MT_ARG_CHECK(capacity >= 0, "mt_intersect_primitives: negative capacity");
    MT_ARG_CHECK(mode >= MT_TRACE_PACKED && mode <= MT_TRACE_PACKED_BINNED,
                 "mt_intersect_primitives: unknown mode %d", mode);
    MT_ARG_CHECK((reinterpret_cast<uintptr_t>(rays) & 15) == 0 && (reinterpret_cast<uintptr_t>(isects) & 15) == 0,
                 "mt_intersect_primitives: rays and isects must be 16-byte aligned");
    if (capacity == 0) return MT_OK;
    TraceArgs A{};
    A.nodes = sc->desc.nodes;
    A.bounds = sc->desc.bounds;
    A.faces = sc->desc.faces;
    A.points = reinterpret_cast<const float4*>(sc->desc.points);
    A.pn = sc->pnodes;
    A.pt = sc->ptris;
    A.root = sc->root_ref;
    A.d_count = d_count;
    A.capacity = capacity;
    const float4* nrm = reinterpret_cast<const float4*>(sc->desc.normals);
    const float2* uvs = reinterpret_cast<const float2*>(sc->desc.uvs);
    float4* R = reinterpret_cast<float4*>(rays);
    float4* I = reinterpret_cast<float4*>(isects);
    dim3 grid(mt_blocks(capacity, 256)), block(256);
    hipStream_t s = ctx->stream;
    // every packed mode gives the same bits; this writer runs one ray per lane
    if (mode == MT_TRACE_REFERENCE) k_intersect_records<1><<<grid, block, 0, s>>>(A, nrm, uvs, R, I);
    else if (mode == MT_TRACE_PACKED_NO_STACK) k_intersect_records<2><<<grid, block, 0, s>>>(A, nrm, uvs, R, I);
    else k_intersect_records<0><<<grid, block, 0, s>>>(A, nrm, uvs, R, I);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_intersect_occlusion(mt_context* ctx, const mt_scene* sc, const void* rays, int32_t* hits,
                                          const int32_t* d_count, int32_t capacity, int32_t mode)
{
    MT_ARG_CHECK(ctx && sc && rays && hits && d_count, "mt_intersect_occlusion: NULL argument");
    MT_ARG_CHECK(capacity >= 0, "mt_intersect_occlusion: negative capacity");
    MT_ARG_CHECK(mode >= MT_TRACE_PACKED && mode <= MT_TRACE_PACKED_BINNED,
                 "mt_intersect_occlusion: unknown mode %d", mode);
    MT_ARG_CHECK((reinterpret_cast<uintptr_t>(rays) & 15) == 0, "mt_intersect_occlusion: rays must be 16-byte aligned");
    if (capacity == 0) return MT_OK;
    TraceArgs A{};
    A.nodes = sc->desc.nodes;
    A.bounds = sc->desc.bounds;
    A.faces = sc->desc.faces;
    A.points = reinterpret_cast<const float4*>(sc->desc.points);
    A.pn = sc->pnodes;
    A.pt = sc->ptris;
    A.root = sc->root_ref;
    A.d_count = d_count;
    A.capacity = capacity;
    const float4* R = reinterpret_cast<const float4*>(rays);
    dim3 grid(mt_blocks(capacity, 256)), block(256);
    hipStream_t s = ctx->stream;
    if (mode == MT_TRACE_REFERENCE) k_occlusion_records<1><<<grid, block, 0, s>>>(A, R, hits);
    else if (mode == MT_TRACE_PACKED_NO_STACK) k_occlusion_records<2><<<grid, block, 0, s>>>(A, R, hits);
    else k_occlusion_records<0><<<grid, block, 0, s>>>(A, R, hits);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_shadow_rays(mt_context* ctx, const float* o_tmin, const float* d_tmax,
                                  const int32_t* prim, const float* t, const float* light_xyz,
                                  const int32_t* d_count, int32_t capacity, float* s_o_tmin,
                                  float* s_d_tmax, int32_t* valid)
{
    MT_ARG_CHECK(ctx && o_tmin && d_tmax && prim && t && light_xyz && d_count && s_o_tmin &&
                     s_d_tmax && valid,
                 "mt_shadow_rays: NULL argument");
    if (capacity <= 0) return MT_OK;
    k_shadow_rays<<<mt_blocks(capacity, 256), 256, 0, ctx->stream>>>(
        reinterpret_cast<const float4*>(o_tmin), reinterpret_cast<const float4*>(d_tmax), prim, t,
        light_xyz[0], light_xyz[1], light_xyz[2], d_count, capacity,
        reinterpret_cast<float4*>(s_o_tmin), reinterpret_cast<float4*>(s_d_tmax), valid);
    MT_LAUNCH_CHECK();
    return MT_OK;
}
