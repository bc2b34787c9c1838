// Per-ray BVH traversal (device code) shared by the traversal entry points
// (mt_trace.hip) and the bounce loop's fused kernels (mt_render.hip):
// intersectMesh / testOcclusion (src/cl/kernel/NormalBVH.cl:14-92,
// 235-317) over the reference arrays and over the packed layout, in the
// reference's visiting order (see mt_trace.hip for the layouts).
#pragma once
#include "mt_devmath.h"
#include "mt_internal.h"

namespace {

struct RayT {
    v3 o, d, inv;
    int sx, sy, sz;
    float tmin, tmax;
};

__device__ __forceinline__ void ray_init(RayT& r, float4 a, float4 b)
{ // initGlobalRay, Util.cl:577-600
    r.o = mk3(a.x, a.y, a.z);
    r.d = mk3(b.x, b.y, b.z);
    r.inv = mk3(1.0f / b.x, 1.0f / b.y, 1.0f / b.z);
    r.sx = r.inv.x < 0.0f;
    r.sy = r.inv.y < 0.0f;
    r.sz = r.inv.z < 0.0f;
    r.tmin = a.w;
    r.tmax = b.w;
}

// intersectBoundT, Util.cl:707-732.  Conditional assignments (not fminf/fmaxf)
// keep the reference's NaN behaviour.  The reference's two early returns
// become terms of one conjunction (all six slab distances are computed
// either way), so the test has no branches: a wave's lanes never split
// inside it, and the node's loads all issue before the first compare.
__device__ __forceinline__ bool box_hit(const RayT& r, float mnx, float mny, float mnz,
                                        float mxx, float mxy, float mxz, float& tnear)
{
    const float tmin = ((r.sx ? mxx : mnx) - r.o.x) * r.inv.x;
    const float tmax = ((r.sx ? mnx : mxx) - r.o.x) * r.inv.x;
    const float tymin = ((r.sy ? mxy : mny) - r.o.y) * r.inv.y;
    const float tymax = ((r.sy ? mny : mxy) - r.o.y) * r.inv.y;
    const float tzmin = ((r.sz ? mxz : mnz) - r.o.z) * r.inv.z;
    const float tzmax = ((r.sz ? mnz : mxz) - r.o.z) * r.inv.z;
    const bool xy = !((tmin > tymax) | (tymin > tmax));
    const float t1 = tymin > tmin ? tymin : tmin;
    const float u1 = tymax < tmax ? tymax : tmax;
    const bool xyz = !((t1 > tzmax) | (tzmin > u1));
    const float t2 = tzmin > t1 ? tzmin : t1;
    const float u2 = (tzmax < u1 ? tzmax : u1) * 1.00000024f;
    tnear = t2;
    return xy & xyz & (t2 < r.tmax) & (u2 > r.tmin);
}

// fastTriangleIntersection, Primitive.cl:3-29, on p1 and the two edges.
// Both early returns of the reference give tMax; here every quantity is
// computed and one select picks t or tMax, so p1 is fetched with the edges
// instead of after the determinant test.
__device__ __forceinline__ float tri_hit(const RayT& r, v3 p1, v3 e1, v3 e2)
{
    const v3 h = cross3(r.d, e2);
    const float a = dot3(e1, h);
    const bool parallel = ((double)a > -0.0000001) & ((double)a < 0.0000001); // double compare, :13
    const float f = 1.0f / a;
    const v3 s = sub3(r.o, p1);
    const float b1 = f * dot3(s, h);
    const v3 q = cross3(s, e1);
    const float b2 = f * dot3(r.d, q);
    const float t = f * dot3(e2, q);
    const bool out = (b1 < 0.0f) | (b1 > 1.0f) | (b2 < 0.0f) | (b1 + b2 > 1.0f) | (t < 0.0f) | (t > r.tmax);
    return (parallel | out) ? r.tmax : t;
}

struct BitStack {
    uint64_t lo, hi;
    __device__ __forceinline__ void push0()
    {
        hi = (hi << 1) | (lo >> 63);
        lo <<= 1;
    }
    __device__ __forceinline__ void pop()
    {
        lo = (lo >> 1) | (hi << 63);
        hi >>= 1;
    }
    __device__ __forceinline__ bool empty() const { return (lo | hi) == 0; }
    // zero bits above the lowest set bit (the pops the reference performs)
    __device__ __forceinline__ int ctz() const
    {
        return lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll(hi);
    }
    __device__ __forceinline__ void shr(int z)
    {
        if (z == 0) return;
        if (z < 64) {
            lo = (lo >> z) | (hi << (64 - z));
            hi >>= z;
        } else {
            lo = hi >> (z - 64);
            hi = 0;
        }
    }
};

// Short stack of postponed far children, newest first.  Each 1 bit of the
// bitstack is a both-children-hit decision whose far child is still to be
// visited; the stack holds the most recent MT_SHORT_STACK of them.  Reaching
// the next pending far child then costs no memory access: the reference's
// pops (walking parent links up to the lowest 1 bit) are counted as
// ctz(bitstack) without being fetched.  When older entries have been
// dropped, the parent walk finds them exactly as the reference does, reading
// parent and sibling links from the current node's record (an interior
// record's last word pair, a leaf triangle's spare words), so no links are
// carried in registers while descending.  Depth N is a trade: each push
// and pop moves N registers, and a deeper stack saves parent walks on deep
// trees.  MT_SHORT_STACK (8) serves the render's 1M-triangle scene; the wave
// pools over sorted batches of deep trees use SHORT_STACK_DEEP.
#ifndef MT_SHORT_STACK
#define MT_SHORT_STACK 8
#endif
template <int N>
struct ShortStackN {
    int ref[N];
    int n;
    __device__ __forceinline__ void push(int r)
    {
#pragma unroll
        for (int k = N - 1; k > 0; k--) ref[k] = ref[k - 1];
        ref[0] = r;
        n = n < N ? n + 1 : N;
    }
    __device__ __forceinline__ int pop()
    {
        const int r = ref[0];
#pragma unroll
        for (int k = 0; k < N - 1; k++) ref[k] = ref[k + 1];
        n--;
        return r;
    }
};
using ShortStack = ShortStackN<MT_SHORT_STACK>;

// -------------------------------------------------------------------------
// Reference layout: NormalBVH.cl:14-92 over nodes[7], bounds[8], faces[10].
// -------------------------------------------------------------------------
template <bool ANY, bool STATS>
__device__ __forceinline__ int trace_reference(const int* __restrict__ nodes,
                                               const float* __restrict__ bounds,
                                               const int* __restrict__ faces,
                                               const float4* __restrict__ points, RayT& r,
                                               int& child, int* st)
{
    int hit = 0;
    int node_id = 0;
    BitStack bs{0, 0};
    int parent_id = 0, sibling_id = 0;
    for (;;) {
        while (!nodes[7 * node_id + 5]) {
            const int* n = nodes + 7 * node_id;
            parent_id = n[4];
            sibling_id = n[1];
            int left = n[2], right = n[3];
            const float* lb = bounds + 8 * nodes[7 * left];
            const float* rb = bounds + 8 * nodes[7 * right];
            float lt, rt;
            bool lh = box_hit(r, lb[0], lb[1], lb[2], lb[4], lb[5], lb[6], lt);
            bool rh = box_hit(r, rb[0], rb[1], rb[2], rb[4], rb[5], rb[6], rt);
            if (STATS) st[0]++;
            if (!lh && !rh) break;
            bs.push0();
            if (lh && rh) {
                node_id = (rt < lt) ? right : left;
                bs.lo |= 1;
            } else {
                node_id = lh ? left : right;
            }
        }
        if (nodes[7 * node_id + 5]) {
            const int* n = nodes + 7 * node_id;
            if (STATS) st[1]++;
            int c = n[6];
            const int* F = faces + 10 * c;
            v3 p1 = xyz(points[F[0]]);
            v3 p2 = xyz(points[F[1]]);
            v3 p3 = xyz(points[F[2]]);
            float t = tri_hit(r, p1, sub3(p2, p1), sub3(p3, p1));
            if ((r.tmin < t) && (t < r.tmax)) {
                if (ANY) return 1;
                hit = 1;
                r.tmax = t;
                child = c;
            }
            parent_id = n[4];
            sibling_id = n[1];
        }
        while ((bs.lo & 1) == 0) {
            if (bs.empty()) return hit;
            node_id = parent_id;
            const int* n = nodes + 7 * node_id;
            parent_id = n[4];
            sibling_id = n[1];
            bs.pop();
            if (STATS) st[2]++;
        }
        node_id = sibling_id;
        bs.lo ^= 1;
    }
}

// -------------------------------------------------------------------------
// Packed layout, same visiting order.  cur >= 0: interior node record;
// cur < 0: leaf ~rank.  parent/sib track the current node's links, which the
// reference reads from the node records (leaves are not records here).
// -------------------------------------------------------------------------
// Resumable traversal state of one ray (packed layout).
// Parent and sibling references of node `cur`: an interior record's last
// word pair, or the spare words of a leaf's triangle record.
__device__ __forceinline__ int2 node_links(const float4* __restrict__ pn, const float4* __restrict__ pt, int cur)
{
    const float4 w = cur >= 0 ? pn[4 * (size_t)cur + 3] : pt[3 * (size_t)(~cur) + 2];
    return make_int2(__float_as_int(w.z), __float_as_int(w.w));
}

// The reference's backtracking (NormalBVH.cl:80-90): climb parent links
// while the lowest bitstack bit is 0, then cross to the sibling and flip the
// bit.  Returns false when the bitstack empties (the ray is finished).
template <bool STATS>
__device__ __forceinline__ bool parent_walk(const float4* __restrict__ pn, const float4* __restrict__ pt, int& cur,
                                            BitStack& bs, int* st)
{
    int2 l = node_links(pn, pt, cur); // (parent, sibling) of cur
    while ((bs.lo & 1) == 0) {
        if (bs.empty()) return false;
        cur = l.x;
        l = node_links(pn, pt, cur);
        bs.pop();
        if (STATS) st[2]++;
    }
    cur = l.y;
    bs.lo ^= 1;
    return true;
}

// Resumable traversal state of one ray (packed layout), short stack of N.
template <int N>
struct TravStateN {
    int cur;
    BitStack bs;
    ShortStackN<N> ss;
    int hit, child;
    int st[3];
    __device__ __forceinline__ void init(int root)
    {
        cur = root;
        bs.lo = 0;
        bs.hi = 0;
        ss.n = 0;
        hit = 0;
        child = -1;
        st[0] = st[1] = st[2] = 0;
    }
};
using TravState = TravStateN<MT_SHORT_STACK>;

// One outer iteration of the packed traversal: descend to a leaf (or a node
// whose children both miss), test the leaf, move to the next pending subtree.
// Returns true when the ray is finished.  The sequence of nodes visited over
// all iterations is exactly the reference's (see trace_packed below).
template <bool ANY, bool STATS, bool SHORT_STACK>
__device__ __forceinline__ bool trace_packed_step(const float4* __restrict__ pn, const float4* __restrict__ pt,
                                                  RayT& r, TravState& S)
{
    while (S.cur >= 0) {
        const float4* N = pn + 4 * (size_t)S.cur;
        float4 a = N[0], b = N[1], c = N[2];
        int4 d = *reinterpret_cast<const int4*>(N + 3);
        float lt, rt;
        bool lh = box_hit(r, a.x, a.y, a.z, a.w, b.x, b.y, lt);
        bool rh = box_hit(r, b.z, b.w, c.x, c.y, c.z, c.w, rt);
        if (STATS) S.st[0]++;
        if (!lh && !rh) break;
        S.bs.push0();
        int nxt;
        if (lh && rh) {
            bool right_first = rt < lt;
            nxt = right_first ? d.y : d.x;
            if (SHORT_STACK) S.ss.push(right_first ? d.x : d.y);
            S.bs.lo |= 1;
        } else {
            nxt = lh ? d.x : d.y;
        }
        S.cur = nxt;
    }
    if (S.cur < 0) {
        if (STATS) S.st[1]++;
        const float4* T = pt + 3 * (size_t)(~S.cur);
        float4 t0 = T[0], t1 = T[1], t2 = T[2];
        float t = tri_hit(r, mk3(t0.x, t0.y, t0.z), mk3(t0.w, t1.x, t1.y), mk3(t1.z, t1.w, t2.x));
        if ((r.tmin < t) && (t < r.tmax)) {
            S.hit = 1;
            if (ANY) return true;
            r.tmax = t;
            S.child = __float_as_int(t2.y);
        }
    }
    if (S.bs.empty()) return true;
    if (SHORT_STACK && S.ss.n > 0) {
        int z = S.bs.ctz();
        if (STATS) S.st[2] += z;
        S.bs.shr(z);
        S.bs.lo ^= 1;
        S.cur = S.ss.pop();
        return false;
    }
    return !parent_walk<STATS>(pn, pt, S.cur, S.bs, S.st);
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{ // set bits of `mask` below this lane (v_mbcnt_lo/hi)
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// One node of the packed traversal per call: an interior node's box pair
// (descending to the chosen child, or moving to the next pending subtree
// when both miss) or a leaf's triangle (then moving on).  Returns true when
// the ray is finished.  Successive calls visit exactly the nodes of
// trace_packed_step in the same order; a wave that calls this in a loop
// keeps every lane on a node of its own ray each iteration instead of
// holding the lanes that reached a leaf until the whole wave has.
template <bool ANY, bool STATS, int DEPTH>
__device__ __forceinline__ bool trace_packed_node(const float4* __restrict__ pn, const float4* __restrict__ pt,
                                                  RayT& r, TravStateN<DEPTH>& S)
{
    if (S.cur >= 0) {
        const float4* N = pn + 4 * (size_t)S.cur;
        float4 a = N[0], b = N[1], c = N[2];
        int4 d = *reinterpret_cast<const int4*>(N + 3);
        float lt, rt;
        bool lh = box_hit(r, a.x, a.y, a.z, a.w, b.x, b.y, lt);
        bool rh = box_hit(r, b.z, b.w, c.x, c.y, c.z, c.w, rt);
        if (STATS) S.st[0]++;
        if (lh || rh) {
            S.bs.push0();
            int nxt;
            if (lh && rh) {
                bool right_first = rt < lt;
                nxt = right_first ? d.y : d.x;
                S.ss.push(right_first ? d.x : d.y);
                S.bs.lo |= 1;
            } else {
                nxt = lh ? d.x : d.y;
            }
            S.cur = nxt;
            return false;
        }
    } else {
        if (STATS) S.st[1]++;
        const float4* T = pt + 3 * (size_t)(~S.cur);
        float4 t0 = T[0], t1 = T[1], t2 = T[2];
        float t = tri_hit(r, mk3(t0.x, t0.y, t0.z), mk3(t0.w, t1.x, t1.y), mk3(t1.z, t1.w, t2.x));
        if ((r.tmin < t) && (t < r.tmax)) {
            S.hit = 1;
            if (ANY) return true;
            r.tmax = t;
            S.child = __float_as_int(t2.y);
        }
    }
    if (S.bs.empty()) return true;
    if (S.ss.n > 0) {
        int z = S.bs.ctz();
        if (STATS) S.st[2] += z;
        S.bs.shr(z);
        S.bs.lo ^= 1;
        S.cur = S.ss.pop();
        return false;
    }
    return !parent_walk<STATS>(pn, pt, S.cur, S.bs, S.st);
}

// Work-groups go to the 8 XCDs in turn (group b on XCD b mod 8), each XCD
// with its own L2.  Where neighbouring work-groups trace rays that walk the
// same subtrees, renumbering the groups so that each XCD takes runs of
// neighbouring work keeps each L2 on its own region's nodes instead of all
// eight caching the same ones, while the XCDs taking the runs in turn keeps
// work that varies across the range balanced between them.
//
// Runs of `chunk` consecutive numbers per XCD (XCD x takes runs x, x + 8,
// x + 16, ...); groups past the last whole round of 8 runs keep their own
// number.  A bijection of [0, groups).
__device__ __forceinline__ unsigned xcd_chunked_group(unsigned b, unsigned groups, unsigned chunk)
{
    const unsigned round = 8u * chunk;
    if (b >= (groups / round) * round) return b;
    const unsigned x = b & 7u, k = b >> 3;
    return ((k / chunk) * 8u + x) * chunk + k % chunk;
}

// Run length of the renumbered kernels (camera rays: 16K paths, a compact
// patch of pixels; binned order: 128K rays, 64 neighbouring origin cells of
// a uniform batch).  Measured: a 1920x1080 x 16 spp render of 1M triangles
// 9.69 -> 9.63 ms with the camera traversal renumbered (runs of 256 or
// 1024: the same); 64M random rays through 4M triangles,
// closest / any hit with the sorted order renumbered 48.4 / 34.0 -> 47.6 /
// 33.1 ms (runs of 16: 47.8 / 33.1; 32: 47.5 / 33.0; 256 or contiguous
// eighths of the order: 48.2 / 33.4-33.6).
constexpr unsigned XCD_RUN = 64;

// Work-group number v of a sweep over a segmented queue: group v takes
// segment v mod nseg from chunk v / nseg (the grid a multiple of nseg).  The
// bounce loop's shading appends work-group b's paths to segment b mod 1024,
// so consecutive segments hold neighbouring pixel blocks; runs of XCD_RUN
// consecutive numbers per XCD keep each XCD's L2 on the nodes of a compact
// stretch of the image (the traversal sweeps: the 1080p x 16 spp render
// 9.56 -> 9.49 ms; the shading sweeps, which read per-path state rather than nodes,
// lost 0.4% with it and keep v = b), and with 1024 segments every group of
// one segment runs on the same XCD.
__device__ __forceinline__ unsigned seg_sweep_group()
{
    return xcd_chunked_group(blockIdx.x, gridDim.x, XCD_RUN);
}

// Idle lanes a pool waits for before starting new rays.
constexpr int POOL_REFILL_MIN = 16;

// Wave-pool traversal of the rays of entries [next, end) (wave-uniform): the
// wave starts one ray per lane, every iteration takes each lane one node
// further on its ray (trace_packed_node), and once POOL_REFILL_MIN lanes'
// rays have ended those lanes take the run's next entries in lane order
// (ballot + mbcnt on a wave-uniform cursor; no atomics), so lanes stay busy
// until the run is drained: a
// wave's tail, where its longest rays run while the other lanes idle, comes
// once per run instead of once per 64 rays.  load(id, r, pn, pt, root)
// initialises entry id's ray (and may set the lane's scene: every lane
// starts on the caller's); done(fin, id, r, S) is called after every
// iteration with the whole wave active, fin set on the lanes whose ray ended
// in it.  Each ray is traversed in the reference's
// visiting order; only the lane that computes it changes.
template <bool ANY, int N = MT_SHORT_STACK, class Load, class Done>
__device__ __forceinline__ void pool_trace(const float4* __restrict__ pn, const float4* __restrict__ pt, int root,
                                           int next, int end, Load&& load, Done&& done)
{
    bool active = false;
    int id = 0;
    RayT r;
    TravStateN<N> S;
    const float4* lpn = pn; // this lane's scene (load may change it)
    const float4* lpt = pt;
    int lroot = root;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        // refill once POOL_REFILL_MIN lanes wait (or all do): starting a ray
        // (two loads, three reciprocals) for one or two lanes at a time would
        // cost the whole wave that block nearly every iteration
        if (idle && next < end && (idle == ~0ull || __popcll(idle) >= POOL_REFILL_MIN)) {
            const int avail = end - next, nidle = (int)__popcll(idle);
            const int n = avail < nidle ? avail : nidle;
            const int rk = (int)lane_rank(idle);
            if (!active && rk < n) {
                id = next + rk;
                load(id, r, lpn, lpt, lroot);
                S.init(lroot);
                active = true;
            }
            next += n;
        }
        if (!__any(active)) break; // run drained, every lane done
        bool fin = false;
        if (active && trace_packed_node<ANY, false>(lpn, lpt, r, S)) {
            fin = true;
            active = false;
        }
        done(fin, id, r, S);
    }
}

// The run of wave w of work-group b in a segmented queue (nseg segments of
// seg_cap entries, counts[p] live in segment p) walked chunk-major with
// runs of `run` entries: group v (seg_sweep_group) takes chunk v / nseg of
// segment v % nseg, 4 runs, one per wave.
__device__ __forceinline__ void pool_seg_run(const int* __restrict__ counts, int nseg, int seg_cap, int run,
                                             int& next, int& end)
{
    const unsigned v = seg_sweep_group();
    const int p = (int)(v % (unsigned)nseg), c = (int)(v / (unsigned)nseg);
    int cnt = counts[p];
    cnt = cnt < seg_cap ? cnt : seg_cap;
    const int off = c * (4 * run) + (int)(threadIdx.x >> 6) * run;
    next = p * seg_cap + off;
    end = p * seg_cap + (cnt < off + run ? cnt : off + run);
}

// Rays per lane of a wave's run in the pool kernels.
constexpr int POOL_RAYS = 4;
// Short-stack depth of the wave pools over sorted (binned) batches: a deep
// tree's rays keep more postponed subtrees in registers (64M random rays
// through 4M triangles, depth 45: see mt_trace.hip).
constexpr int SHORT_STACK_DEEP = 12;

template <bool ANY, bool STATS, bool SHORT_STACK>
__device__ __forceinline__ int trace_packed(const float4* __restrict__ pn,
                                            const float4* __restrict__ pt, int root, RayT& r,
                                            int& child, int* st)
{
    int hit = 0;
    int cur = root;
    BitStack bs{0, 0};
    ShortStack ss;
    ss.n = 0;
    for (;;) {
        while (cur >= 0) {
            const float4* N = pn + 4 * (size_t)cur;
            float4 a = N[0], b = N[1], c = N[2];
            int4 d = *reinterpret_cast<const int4*>(N + 3);
            float lt, rt;
            bool lh = box_hit(r, a.x, a.y, a.z, a.w, b.x, b.y, lt);
            bool rh = box_hit(r, b.z, b.w, c.x, c.y, c.z, c.w, rt);
            if (STATS) st[0]++;
            if (!lh && !rh) break;
            bs.push0();
            int nxt;
            if (lh && rh) {
                bool right_first = rt < lt;
                nxt = right_first ? d.y : d.x;
                if (SHORT_STACK) ss.push(right_first ? d.x : d.y);
                bs.lo |= 1;
            } else {
                nxt = lh ? d.x : d.y;
            }
            cur = nxt;
        }
        if (cur < 0) {
            if (STATS) st[1]++;
            const float4* T = pt + 3 * (size_t)(~cur);
            float4 t0 = T[0], t1 = T[1], t2 = T[2];
            float t = tri_hit(r, mk3(t0.x, t0.y, t0.z), mk3(t0.w, t1.x, t1.y), mk3(t1.z, t1.w, t2.x));
            if ((r.tmin < t) && (t < r.tmax)) {
                if (ANY) return 1;
                hit = 1;
                r.tmax = t;
                child = __float_as_int(t2.y);
            }
        }
        if (bs.empty()) return hit;
        if (SHORT_STACK && ss.n > 0) {
            // the lowest 1 bit's far child is on top of the short stack
            int z = bs.ctz();
            if (STATS) st[2] += z;
            bs.shr(z);
            bs.lo ^= 1;
            cur = ss.pop();
            continue;
        }
        if (!parent_walk<STATS>(pn, pt, cur, bs, st)) return hit;
    }
}

} // namespace
