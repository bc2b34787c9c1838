// Device BVH builder: CNormalBVH.build/subdivide (src/cl/scene/
// CNormalBVH.java:45-148) on the GPU, producing the same nodes and bounds as
// the host builder mt_bvh_build (same node numbering, same primitive order
// in the leaves, bounds equal up to the sign of zero).
//
// The reference recurses depth first; here the tree grows one level per
// round and every node of a level is processed at once:
//   1. bounds: each element (a position in the object order) adds its
//      primitive box and centroid to its node's bounds — wave-segmented
//      min/max over the runs of equal node id (positions of one node are
//      contiguous), then one atomic per run on order-preserving integer
//      keys;
//   2. per node: write bounds and the node record; a single primitive makes
//      a leaf; otherwise the split axis (largest centroid extent) and the
//      centroid midpoint;
//   3. flags: right = centroid[axis] >= split, counted per node;
//   4. one exclusive scan of the right flags over all positions;
//   5. per node: mid = start + lefts (or the midpoint index when one side is
//      empty), child numbering from the depth-first counter formula (left
//      child `base`, right `base + 1`, left subtree from `base + 2`, right
//      subtree from `base + 2 * lefts`), child records;
//   6. the two-pointer partition's result, computed in parallel: the k-th
//      right element from the left of [start, mid) trades places with the
//      k-th left element from the right of [mid, end) — exactly the swaps
//      the sequential partition performs;
//   7. each element takes its child's id for the next level.
#include "mt_internal.h"

namespace {

constexpr int BB = 256;

__device__ __forceinline__ int fkey(float f)
{ // order-preserving int key of a float (min/max on keys = min/max on floats)
    int b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float fval(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }

struct BuildState {
    int n;
    float4 *pmn, *pmx, *cen;       // per primitive
    int* obj;                      // object order (current)
    int* obj_next;
    int* seg;                      // node id of each position, -1 once in a finished leaf
    int* flag;                     // right flag per position (0 for inactive)
    int* scan;                     // exclusive scan of flag, [n] = total
    int* table;                    // pairing table, indexed start + k
    // per node
    int *nstart, *nend, *nbase, *nmid, *nkind; // kind: 0 interior, 1 leaf, 2 interior without swaps
    int* naxis;
    float* nsplit;
    int* key;                      // 12 keys per node: bb min xyz, bb max xyz, bc min xyz, bc max xyz
    int* nleft;                    // left count per node
    int* nodes;
    float* bounds;
};

__global__ void k_prims(BuildState S, const float4* __restrict__ points, const int* __restrict__ faces)
{ // CMesh.getBound (CMesh.java:64-71) and the box centre per primitive
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; k++) {
        float4 p = points[faces[10 * (size_t)i + k]];
        float c[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; a++) { // CBound.include, CBound.java:54-70
            if (c[a] < mn[a]) mn[a] = c[a];
            if (c[a] > mx[a]) mx[a] = c[a];
        }
    }
    S.pmn[i] = make_float4(mn[0], mn[1], mn[2], 0.0f);
    S.pmx[i] = make_float4(mx[0], mx[1], mx[2], 0.0f);
    S.cen[i] = make_float4(0.5f * (mn[0] + mx[0]), 0.5f * (mn[1] + mx[1]), 0.5f * (mn[2] + mx[2]), 0.0f);
    S.obj[i] = i;
    S.seg[i] = 0;
}

__global__ void k_root(BuildState S, int* list0, int* cnt)
{
    if (threadIdx.x == 0) {
        S.nstart[0] = 0;
        S.nend[0] = S.n;
        S.nbase[0] = 1;
        for (int k = 0; k < 7; k++) S.nodes[k] = 0;
        list0[0] = 0;
        cnt[1] = S.n; // scan length
    }
}

__global__ void k_zero1(int* p)
{
    if (threadIdx.x == 0) *p = 0;
}

__global__ void k_level_init(BuildState S, const int* __restrict__ list, int count)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int s = list[t];
    int* K = S.key + 12 * (size_t)s;
    for (int a = 0; a < 3; a++) {
        K[a] = fkey(INFINITY);
        K[3 + a] = fkey(-INFINITY);
        K[6 + a] = fkey(INFINITY);
        K[9 + a] = fkey(-INFINITY);
    }
    S.nleft[s] = 0;
}

// Segmented inclusive scan over the wave's runs of equal `s`.
template <class Op>
__device__ __forceinline__ int seg_scan(int v, int s, int lane, Op op)
{
    for (int o = 1; o < 64; o <<= 1) {
        int ov = __shfl_up(v, o), os = __shfl_up(s, o);
        if (lane >= o && os == s) v = op(v, ov);
    }
    return v;
}

// Last lane of its run; call with the whole wave active (a shuffle reads
// nothing useful from an inactive lane).
__device__ __forceinline__ bool run_tail(int s, int lane)
{
    int ns = __shfl_down(s, 1);
    return lane == 63 || ns != s;
}

// Each wave reduces BOUND_ROWS consecutive rows of 64 positions.  A run of
// equal node id still open at the end of a row is carried into the next row
// instead of being flushed, so a node gets one set of atomics per run per
// wave, not per 64 positions: near the root, where one node spans every
// position, 16x fewer atomics land on the same addresses.  Min/max and sums
// do not depend on the order, so results are unchanged.
constexpr int BOUND_ROWS = 16;

// Lanes [0, result) of the row hold its first run.
__device__ __forceinline__ int first_change(int s, int lane)
{
    const int prev = __shfl_up(s, 1);
    const uint64_t ch = __ballot(lane > 0 && prev != s);
    return ch ? __builtin_ctzll(ch) : 64;
}

// Keys 0-2 and 6-8 are minima, 3-5 and 9-11 maxima.
__device__ __forceinline__ int kcomb(int k, int x, int y)
{
    return ((k / 3) & 1) ? (x > y ? x : y) : (x < y ? x : y);
}

__device__ __forceinline__ void flush_keys(const BuildState& S, int s, const int (&v)[12])
{
    int* K = S.key + 12 * (size_t)s;
    for (int k = 0; k < 3; k++) atomicMin(&K[k], v[k]);
    for (int k = 3; k < 6; k++) atomicMax(&K[k], v[k]);
    for (int k = 6; k < 9; k++) atomicMin(&K[k], v[k]);
    for (int k = 9; k < 12; k++) atomicMax(&K[k], v[k]);
}

__global__ __launch_bounds__(BB) void k_bounds(BuildState S)
{
    const int lane = __lane_id();
    const int64_t base = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * BOUND_ROWS;
    if (base >= S.n) return; // wave-uniform
    auto mn = [](int x, int y) { return x < y ? x : y; };
    auto mx = [](int x, int y) { return x > y ? x : y; };
    int cs = -1; // node of the run carried from the previous row (wave-uniform)
    int cv[12];
    for (int r = 0; r < BOUND_ROWS; r++) {
        const int64_t row = base + (int64_t)r * 64;
        if (row >= S.n) break; // wave-uniform
        const int64_t i = row + lane;
        int s = -1;
        float4 a = make_float4(0, 0, 0, 0), b = a, c = a;
        if (i < S.n) {
            s = S.seg[i];
            if (s >= 0) {
                int p = S.obj[i];
                a = S.pmn[p];
                b = S.pmx[p];
                c = S.cen[p];
            }
        }
        int v[12];
        v[0] = mn(fkey(a.x), fkey(b.x));
        v[1] = mn(fkey(a.y), fkey(b.y));
        v[2] = mn(fkey(a.z), fkey(b.z));
        v[3] = mx(fkey(a.x), fkey(b.x));
        v[4] = mx(fkey(a.y), fkey(b.y));
        v[5] = mx(fkey(a.z), fkey(b.z));
        v[6] = fkey(c.x);
        v[7] = fkey(c.y);
        v[8] = fkey(c.z);
        v[9] = v[6];
        v[10] = v[7];
        v[11] = v[8];
#pragma unroll
        for (int k = 0; k < 3; k++) v[k] = seg_scan(v[k], s, lane, mn);
#pragma unroll
        for (int k = 3; k < 6; k++) v[k] = seg_scan(v[k], s, lane, mx);
#pragma unroll
        for (int k = 6; k < 9; k++) v[k] = seg_scan(v[k], s, lane, mn);
#pragma unroll
        for (int k = 9; k < 12; k++) v[k] = seg_scan(v[k], s, lane, mx);
        const int fc = first_change(s, lane);
        const int s0 = __shfl(s, 0);
        if (cs >= 0) {
            if (s0 == cs) {
                if (lane < fc) {
#pragma unroll
                    for (int k = 0; k < 12; k++) v[k] = kcomb(k, v[k], cv[k]);
                }
            } else if (lane == 0) {
                flush_keys(S, cs, cv);
            }
        }
        const bool tail = run_tail(s, lane); // all lanes take part in the shuffle
        if (s >= 0 && tail && lane != 63) flush_keys(S, s, v);
        cs = __shfl(s, 63);
#pragma unroll
        for (int k = 0; k < 12; k++) cv[k] = __shfl(v[k], 63);
    }
    if (cs >= 0 && lane == 0) flush_keys(S, cs, cv);
}

__global__ void k_node_a(BuildState S, const int* __restrict__ list, int count)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int s = list[t];
    const int* K = S.key + 12 * (size_t)s;
    float* pb = S.bounds + 8 * (size_t)s;
    pb[0] = fval(K[0]);
    pb[1] = fval(K[1]);
    pb[2] = fval(K[2]);
    pb[3] = 0.0f;
    pb[4] = fval(K[3]);
    pb[5] = fval(K[4]);
    pb[6] = fval(K[5]);
    pb[7] = 0.0f;
    int* pn = S.nodes + 7 * (size_t)s;
    pn[0] = s;
    int start = S.nstart[s], end = S.nend[s];
    if (end - start < 2) {
        pn[6] = S.obj[start];
        pn[5] = 1;
        S.nkind[s] = 1;
        return;
    }
    S.nkind[s] = 0;
    float cmn[3] = {fval(K[6]), fval(K[7]), fval(K[8])}, cmx[3] = {fval(K[9]), fval(K[10]), fval(K[11])};
    float dx = cmx[0] - cmn[0], dy = cmx[1] - cmn[1], dz = cmx[2] - cmn[2];
    int dim = (dx > dy && dx > dz) ? 0 : (dy > dz ? 1 : 2);
    S.naxis[s] = dim;
    S.nsplit[s] = 0.5f * (cmn[dim] + cmx[dim]);
}

__global__ __launch_bounds__(BB) void k_flags(BuildState S)
{ // rows and carried runs as in k_bounds; the left count per node is a sum
    const int lane = __lane_id();
    const int64_t base = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * BOUND_ROWS;
    if (base >= S.n) return; // wave-uniform
    auto add = [](int x, int y) { return x + y; };
    int cs = -1, cv = 0;
    for (int r = 0; r < BOUND_ROWS; r++) {
        const int64_t row = base + (int64_t)r * 64;
        if (row >= S.n) break; // wave-uniform
        const int64_t i = row + lane;
        int s = -1, right = 0;
        if (i < S.n) {
            s = S.seg[i];
            if (s >= 0 && S.nkind[s] == 0) {
                const float4 c = S.cen[S.obj[i]];
                const int dim = S.naxis[s];
                const float v = dim == 0 ? c.x : (dim == 1 ? c.y : c.z);
                right = !(v < S.nsplit[s]) ? 1 : 0; // the partition keeps centroid < split on the left
            } else {
                s = -1;
            }
            S.flag[i] = right;
        }
        int left = (s >= 0 && !right) ? 1 : 0;
        left = seg_scan(left, s, lane, add);
        const int fc = first_change(s, lane);
        const int s0 = __shfl(s, 0);
        if (cs >= 0) {
            if (s0 == cs) {
                if (lane < fc) left += cv;
            } else if (lane == 0 && cv) {
                atomicAdd(&S.nleft[cs], cv);
            }
        }
        const bool tail = run_tail(s, lane); // all lanes take part in the shuffle
        if (s >= 0 && tail && lane != 63 && left) atomicAdd(&S.nleft[s], left);
        cs = __shfl(s, 63);
        cv = __shfl(left, 63);
    }
    if (cs >= 0 && lane == 0 && cv) atomicAdd(&S.nleft[cs], cv);
}

// The children of the level's split nodes are appended to the next level's
// list with one atomic per wave (the list's order does not matter: every
// node's numbering comes from its range, not from its place in the list).
__global__ __launch_bounds__(BB) void k_node_b(BuildState S, const int* __restrict__ list, int count,
                                               int* __restrict__ next, int* __restrict__ next_count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)blockIdx.x * blockDim.x >= count) return; // whole work-group idle
    int s = -1;
    bool split = false;
    if (t < count) {
        s = list[t];
        split = S.nkind[s] != 1;
    }
    int li = 0, ri = 0;
    if (split) {
        int start = S.nstart[s], end = S.nend[s], base = S.nbase[s];
        int nl = S.nleft[s];
        int mid = start + nl;
        if (mid == start || mid == end) { // one side empty: no swaps, split the range in half
            mid = start + (end - start) / 2;
            S.nkind[s] = 2;
        }
        S.nmid[s] = mid;
        li = base;
        ri = base + 1;
        int* pn = S.nodes + 7 * (size_t)s;
        pn[2] = li;
        pn[3] = ri;
        int* ln = S.nodes + 7 * (size_t)li;
        int* rn = S.nodes + 7 * (size_t)ri;
        for (int k = 0; k < 7; k++) {
            ln[k] = 0;
            rn[k] = 0;
        }
        ln[4] = s;
        rn[4] = s;
        ln[1] = ri;
        rn[1] = li;
        S.nstart[li] = start;
        S.nend[li] = mid;
        S.nbase[li] = base + 2;
        S.nstart[ri] = mid;
        S.nend[ri] = end;
        S.nbase[ri] = base + 2 * (mid - start);
    }
    const uint64_t m = __ballot(split); // whole wave active here
    int wbase = 0;
    if (m) {
        if (__lane_id() == 0) wbase = atomicAdd(next_count, 2 * (int)__popcll(m));
        wbase = __shfl(wbase, 0);
    }
    if (split) {
        const int w = wbase + 2 * (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        next[w] = li;
        next[w + 1] = ri;
    }
}

// Right part [mid, end): the k-th left element counted from the right
// registers its position at table[start + k].
__global__ __launch_bounds__(BB) void k_pair(BuildState S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    const int s = S.seg[i];
    if (s < 0 || S.nkind[s] != 0) return;
    const int mid = S.nmid[s];
    if (i < mid || S.flag[i]) return;
    const int end = S.nend[s];
    const int rights_after = S.scan[end] - S.scan[i + 1];
    const int k = (end - 1 - i) - rights_after; // left elements in (i, end)
    S.table[S.nstart[s] + k] = i;
}

// New object order and next-level node ids.
__global__ __launch_bounds__(BB) void k_scatter(BuildState S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    const int s = S.seg[i];
    if (s < 0) {
        S.obj_next[i] = S.obj[i];
        return;
    }
    const int kind = S.nkind[s];
    if (kind == 1) { // leaf: done
        S.obj_next[i] = S.obj[i];
        S.seg[i] = -1;
        return;
    }
    const int mid = S.nmid[s];
    const int start = S.nstart[s];
    if (kind == 0 && i < mid && S.flag[i]) {
        const int k = S.scan[i] - S.scan[start]; // right elements in [start, i)
        const int q = S.table[start + k];
        S.obj_next[i] = S.obj[q];
        S.obj_next[q] = S.obj[i];
    } else if (!(kind == 0 && i >= mid && !S.flag[i])) {
        S.obj_next[i] = S.obj[i];
    }
    S.seg[i] = i < mid ? S.nbase[s] : S.nbase[s] + 1;
}

} // namespace

extern "C" int32_t mt_bvh_build_device(mt_context* ctx, int32_t ntris, const float* points, const int32_t* faces,
                                       int32_t* nodes, float* bounds, int32_t* max_depth)
{
    MT_ARG_CHECK(ctx && points && faces && nodes && bounds, "mt_bvh_build_device: NULL argument");
    MT_ARG_CHECK(ntris >= 1 && ntris <= (1 << 30), "mt_bvh_build_device: ntris %d", ntris);
    const size_t n = (size_t)ntris, m = 2 * n - 1;
    size_t bytes = 0;
    auto take = [&](size_t b) { size_t off = bytes; bytes += (b + 255) & ~(size_t)255; return off; };
    size_t oMn = take(16 * n), oMx = take(16 * n), oCen = take(16 * n);
    size_t oObj = take(4 * n), oObj2 = take(4 * n), oSeg = take(4 * n), oFlag = take(4 * n), oScan = take(4 * (n + 1));
    size_t oTab = take(4 * n);
    size_t oStart = take(4 * m), oEnd = take(4 * m), oBase = take(4 * m), oMid = take(4 * m), oKind = take(4 * m);
    size_t oAxis = take(4 * m), oSplit = take(4 * m), oKey = take(48 * m), oLeft = take(4 * m);
    size_t oList0 = take(4 * m), oList1 = take(4 * m), oCnt = take(4 * 4);
    char* mem = nullptr;
    if (hipMalloc(&mem, bytes) != hipSuccess) {
        mt_set_error("mt_bvh_build_device: cannot allocate %.2f GB of build state", bytes / 1e9);
        return MT_ERR_ALLOC;
    }
    auto I = [&](size_t o) { return reinterpret_cast<int*>(mem + o); };
    BuildState S;
    S.n = ntris;
    S.pmn = reinterpret_cast<float4*>(mem + oMn);
    S.pmx = reinterpret_cast<float4*>(mem + oMx);
    S.cen = reinterpret_cast<float4*>(mem + oCen);
    S.obj = I(oObj);
    S.obj_next = I(oObj2);
    S.seg = I(oSeg);
    S.flag = I(oFlag);
    S.scan = I(oScan);
    S.table = I(oTab);
    S.nstart = I(oStart);
    S.nend = I(oEnd);
    S.nbase = I(oBase);
    S.nmid = I(oMid);
    S.nkind = I(oKind);
    S.naxis = I(oAxis);
    S.nsplit = reinterpret_cast<float*>(mem + oSplit);
    S.key = I(oKey);
    S.nleft = I(oLeft);
    S.nodes = nodes;
    S.bounds = bounds;
    int* list[2] = {I(oList0), I(oList1)};
    int* cnt = I(oCnt);
    hipStream_t st = ctx->stream;
    int32_t rc = MT_OK;
    const int grid = mt_blocks((int64_t)n, BB);
    // Build state per level; hipMemcpy below synchronises the stream.
    k_prims<<<grid, BB, 0, st>>>(S, reinterpret_cast<const float4*>(points), faces);
    k_root<<<1, 64, 0, st>>>(S, list[0], cnt);
    int count = 1, level = 0, cur = 0;
    while (rc == MT_OK && count > 0) {
        const int lg = mt_blocks(count, BB);
        k_level_init<<<lg, BB, 0, st>>>(S, list[cur], count);
        k_bounds<<<mt_blocks(n, BB * BOUND_ROWS), BB, 0, st>>>(S);
        k_node_a<<<lg, BB, 0, st>>>(S, list[cur], count);
        k_flags<<<mt_blocks(n, BB * BOUND_ROWS), BB, 0, st>>>(S);
        rc = mt_scan_excl_i32(ctx, S.flag, S.scan, cnt + 1, ntris, S.scan + n);
        if (rc) break;
        k_zero1<<<1, 64, 0, st>>>(cnt);
        k_node_b<<<lg, BB, 0, st>>>(S, list[cur], count, list[1 - cur], cnt);
        k_pair<<<grid, BB, 0, st>>>(S);
        k_scatter<<<grid, BB, 0, st>>>(S);
        if (hipGetLastError() != hipSuccess) {
            mt_set_error("mt_bvh_build_device: launch failed");
            rc = MT_ERR_HIP;
            break;
        }
        std::swap(S.obj, S.obj_next);
        int next = 0;
        if (hipMemcpyAsync(&next, cnt, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            mt_set_error("mt_bvh_build_device: level readback failed");
            rc = MT_ERR_HIP;
            break;
        }
        if (next > 0) level++;
        count = next;
        cur = 1 - cur;
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(mem);
    if (rc == MT_OK && max_depth) *max_depth = level;
    if (rc == MT_OK) rc = mt_scan_check(ctx);
    return rc;
}
