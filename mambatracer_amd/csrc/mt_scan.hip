// Stage D: single-pass exclusive scan and stable stream compaction.
//
// Replaces PrefixSumInteger's four launches + host readback
// (src/cl/algorithms/PrefixSumInteger.java:80-98; kernels Util.cl:917-980)
// and CCompact's seven launches (src/cl/algorithms/CCompact.java:79-89;
// CompactAtomic, Compact.cl:21-40) with one kernel each.
//
// Each work-group takes a tile of 8192 elements (256 lanes x 32 rows, element
// base + row*256 + lane, so every load is a coalesced 1 KB wave access).
// Large tiles keep 32 loads per lane in flight and cut the tickets and
// look-backs per element: on MI355X, 2048-element tiles moved a 33M-slot
// render queue at 4.6 TB/s, 8192-element tiles at 5.2 TB/s (a plain
// device copy: 5.2 TB/s).
// Within a tile, 0/1 predicates are ranked with wave64 ballots and
// mbcnt (no LDS scan at all); integer values use a 6-step shuffle scan per
// wave.  Tiles are chained by decoupled look-back: a tile publishes its
// aggregate, then wave 0 inspects up to 64 predecessors at once until it
// meets an inclusive prefix.  Tile numbers come from an atomic ticket, so a
// tile only ever waits on tiles that started before it; the wait is also
// bounded by a spin limit that raises an error flag instead of hanging.
// Status words carry a per-call epoch and never need clearing.  The epoch
// lives on the device (advanced by the reset kernel that precedes every
// scan), so a captured graph that replays a scan draws a fresh epoch on
// every replay.
//
// Deviation from the reference: localScanInteger's exclusive conversion
// `aux[l] > 0 ? aux[l-1] : aux[0]` (Util.cl:941-944) reads aux[-1] for lane 0
// and is exclusive only for non-negative inputs; this scan is the exclusive
// prefix sum for any int32 input (wrapping), which is what the reference's
// callers expect (PrefixSumIntegerTest.java:30-36, LightConfiguration.cl:
// 31-55).  CompactAtomic's order is whatever its atomic_inc produces; this
// compaction is stable (survivors in index order), matching prepareLightInfo.
#include "mt_internal.h"

#include <mutex>

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ROWS = 32;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ROWS;
constexpr int SCAN_WAVES = SCAN_THREADS / 64;
constexpr uint32_t FLAG_AGG = 1u, FLAG_INC = 2u;
constexpr uint32_t SPIN_LIMIT = 1u << 26;

struct ScanWork {
    uint32_t* ticket; // [0] ticket, [1] error flag (spin limit hit), [2] epoch of the current call
    uint64_t* status; // [ntiles]
};

// Tile I/O goes through buffer resources ("windows"): a descriptor whose
// range check drops stores and zeroes loads at byte offsets at or past its
// size.  Tails and lanes with nothing to write then need no branches, so
// the compiler keeps all of a tile's loads in flight together instead of
// serialising one guarded load per row.
constexpr int RSRC_WORD3 = 0x00020000; // raw (untyped) buffer, 32-bit data format
constexpr uint32_t OUT_OF_RANGE = 0x7fffff00u;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t window(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, RSRC_WORD3);
}

template <int B> struct BufIO;
template <> struct BufIO<4> {
    typedef unsigned int T;
    static __device__ T ld(__amdgpu_buffer_rsrc_t r, uint32_t o) { return __builtin_amdgcn_raw_buffer_load_b32(r, o, 0, 0); }
    static __device__ void st(T v, __amdgpu_buffer_rsrc_t r, uint32_t o) { __builtin_amdgcn_raw_buffer_store_b32(v, r, o, 0, 0); }
};
template <> struct BufIO<8> {
    typedef u32x2 T;
    static __device__ T ld(__amdgpu_buffer_rsrc_t r, uint32_t o) { return __builtin_amdgcn_raw_buffer_load_b64(r, o, 0, 0); }
    static __device__ void st(T v, __amdgpu_buffer_rsrc_t r, uint32_t o) { __builtin_amdgcn_raw_buffer_store_b64(v, r, o, 0, 0); }
};
template <> struct BufIO<16> {
    typedef u32x4 T;
    static __device__ T ld(__amdgpu_buffer_rsrc_t r, uint32_t o) { return __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, 0); }
    static __device__ void st(T v, __amdgpu_buffer_rsrc_t r, uint32_t o) { __builtin_amdgcn_raw_buffer_store_b128(v, r, o, 0, 0); }
};

__device__ __forceinline__ uint64_t pack_status(uint32_t epoch, uint32_t flag, uint32_t v)
{
    return ((uint64_t)(epoch & 0x3FFFFFFFu) << 34) | ((uint64_t)flag << 32) | (uint64_t)v;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{ // number of set bits of `mask` below this lane (v_mbcnt_lo/hi)
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Decoupled look-back by wave 0 of the tile.  Returns the exclusive prefix of
// `tile` (all lanes of wave 0 agree); publishes the tile's inclusive prefix.
__device__ uint32_t look_back(ScanWork W, int tile, uint32_t aggregate, uint32_t epoch)
{
    const int lane = lane_id();
    uint64_t* st = W.status;
    if (tile == 0) {
        if (lane == 0)
            __hip_atomic_store(&st[0], pack_status(epoch, FLAG_INC, aggregate), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return 0u;
    }
    if (lane == 0)
        __hip_atomic_store(&st[tile], pack_status(epoch, FLAG_AGG, aggregate), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    uint32_t prefix = 0;
    int window_end = tile - 1; // predecessor examined by lane 0
    const uint32_t ep = epoch & 0x3FFFFFFFu;
    uint32_t spins = 0;
    for (;;) {
        int idx = window_end - lane;
        uint64_t w;
        uint32_t flag, val;
        for (;;) {
            if (idx >= 0) {
                w = __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bool cur = (uint32_t)(w >> 34) == ep;
                flag = cur ? (uint32_t)(w >> 32) & 3u : 0u;
                val = (uint32_t)w;
            } else { // before tile 0: an inclusive zero
                flag = FLAG_INC;
                val = 0;
            }
            if (__all(flag != 0)) break;
            if (++spins > SPIN_LIMIT) {
                if (lane == 0) atomicOr(&W.ticket[1], 1u);
                return prefix; // results invalid; the host reports the error flag
            }
            __builtin_amdgcn_s_sleep(1);
        }
        uint64_t inc = __ballot(flag == FLAG_INC);
        // lanes up to and including the nearest inclusive predecessor contribute
        int first_inc = inc ? __builtin_ctzll(inc) : 64;
        uint32_t contrib = lane <= first_inc ? val : 0u;
        for (int o = 32; o > 0; o >>= 1) contrib += __shfl_xor(contrib, o);
        prefix += contrib;
        if (inc) break;
        window_end -= 64;
    }
    if (lane == 0)
        __hip_atomic_store(&st[tile], pack_status(epoch, FLAG_INC, prefix + aggregate),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return prefix;
}

__global__ void k_scan_reset(ScanWork W, const int* __restrict__ d_len, int capacity,
                             int* __restrict__ d_total)
{
    if (threadIdx.x == 0) {
        W.ticket[0] = 0u;
        uint32_t e = (W.ticket[2] + 1u) & 0x3FFFFFFFu; // epoch 0 is never current
        W.ticket[2] = e ? e : 1u;
        int len = *d_len;
        if (len > capacity) len = capacity;
        if (len <= 0 && d_total) *d_total = 0;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan(const int* __restrict__ in,
                                                       int* __restrict__ out,
                                                       const int* __restrict__ d_len,
                                                       int capacity, int* __restrict__ d_total,
                                                       ScanWork W)
{
    __shared__ int s_tile;
    __shared__ uint32_t s_wave[SCAN_ROWS][SCAN_WAVES];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t epoch = W.ticket[2];
    int len = *d_len;
    if (len > capacity) len = capacity;
    // Work-groups beyond the tiles the live length needs leave without a
    // ticket (so the ticket counter is not serialised over empty groups);
    // the remaining groups draw exactly tiles 0 .. ntiles-1.
    if ((int64_t)blockIdx.x * SCAN_TILE >= len) return;
    if (tid == 0) s_tile = (int)atomicAdd(&W.ticket[0], 1u);
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    const int64_t base = (int64_t)tile * SCAN_TILE;
    const int64_t rest = (int64_t)len - base;
    const uint32_t live = (uint32_t)(rest < SCAN_TILE ? rest : SCAN_TILE);
    // Loads past the live length read 0 (buffer range check), the identity.
    const __amdgpu_buffer_rsrc_t rin = window(in + base, live * 4);
    uint32_t v[SCAN_ROWS], ex[SCAN_ROWS];
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++) v[k] = BufIO<4>::ld(rin, (uint32_t)(k * SCAN_THREADS + tid) * 4);
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++) {
        uint32_t x = v[k];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        ex[k] = x - v[k];
        if (lane == 63) s_wave[k][wave] = x;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int k = 0; k < SCAN_ROWS; k++)
            for (int w = 0; w < SCAN_WAVES; w++) {
                uint32_t t = s_wave[k][w];
                s_wave[k][w] = run;
                run += t;
            }
        s_prefix = run; // tile aggregate, replaced by the exclusive prefix below
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t agg = s_prefix;
        uint32_t pre = look_back(W, tile, agg, epoch);
        if (lane == 0) {
            s_prefix = pre;
            if ((int64_t)len - 1 < base + SCAN_TILE && d_total) *d_total = (int)(pre + agg);
        }
    }
    __syncthreads();
    const uint32_t pre = s_prefix;
    const __amdgpu_buffer_rsrc_t rout = window(out + base, live * 4);
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++)
        BufIO<4>::st(pre + s_wave[k][wave] + ex[k], rout, (uint32_t)(k * SCAN_THREADS + tid) * 4);
}

struct PayloadSet {
    const void* src[8];
    void* dst[8];
    int bytes[8];
    int n;
};

// Payload moves go through buffer windows: a tile's source window
// [base, base+live) and destination window [pre, pre+agg).  Lanes without a
// survivor store to an out-of-range offset, so every load and store is
// unconditional: no branches, all of a stream's loads in flight before its
// first store.
// dst_off[k]: byte offset of row k's survivor in the destination window
// divided by the element size, or OUT_OF_RANGE for a lane without one.
template <int B>
__device__ __forceinline__ void move_stream(const void* src, void* dst, int64_t base, uint32_t live,
                                            uint32_t pre, uint32_t agg, const uint32_t (&dst_off)[SCAN_ROWS],
                                            int tid)
{
    const __amdgpu_buffer_rsrc_t rs = window(static_cast<const char*>(src) + base * B, live * B);
    const __amdgpu_buffer_rsrc_t rd = window(static_cast<char*>(dst) + (int64_t)pre * B, agg * B);
    typename BufIO<B>::T v[SCAN_ROWS];
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++) v[k] = BufIO<B>::ld(rs, (uint32_t)(k * SCAN_THREADS + tid) * B);
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++)
        BufIO<B>::st(v[k], rd, dst_off[k] == OUT_OF_RANGE ? OUT_OF_RANGE : dst_off[k] * B);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_compact(const int* __restrict__ flags,
                                                          int flag_mode,
                                                          const int* __restrict__ d_len,
                                                          int capacity,
                                                          int* __restrict__ out_index,
                                                          int* __restrict__ d_count,
                                                          PayloadSet P, ScanWork W)
{
    __shared__ int s_tile;
    __shared__ uint32_t s_wave[SCAN_ROWS][SCAN_WAVES];
    __shared__ uint32_t s_prefix, s_agg;
    const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t epoch = W.ticket[2];
    int len = *d_len;
    if (len > capacity) len = capacity;
    // Work-groups beyond the tiles the live length needs leave without a
    // ticket (so the ticket counter is not serialised over empty groups);
    // the remaining groups draw exactly tiles 0 .. ntiles-1.
    if ((int64_t)blockIdx.x * SCAN_TILE >= len) return;
    if (tid == 0) s_tile = (int)atomicAdd(&W.ticket[0], 1u);
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    const int64_t base = (int64_t)tile * SCAN_TILE;
    const int64_t rest = (int64_t)len - base;
    const uint32_t live = (uint32_t)(rest < SCAN_TILE ? rest : SCAN_TILE);
    const __amdgpu_buffer_rsrc_t rflag = window(flags + base, live * 4);
    int x[SCAN_ROWS];
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++) x[k] = (int)BufIO<4>::ld(rflag, (uint32_t)(k * SCAN_THREADS + tid) * 4);
    uint64_t mask[SCAN_ROWS];
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++) {
        const bool f = (uint32_t)(k * SCAN_THREADS + tid) < live && (flag_mode == 0 ? (x[k] != 0) : (x[k] >= 0));
        mask[k] = __ballot(f);
        if (lane == 0) s_wave[k][wave] = (uint32_t)__popcll(mask[k]);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int k = 0; k < SCAN_ROWS; k++)
            for (int w = 0; w < SCAN_WAVES; w++) {
                uint32_t t = s_wave[k][w];
                s_wave[k][w] = run;
                run += t;
            }
        s_prefix = run;
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t agg = s_prefix;
        uint32_t pre = look_back(W, tile, agg, epoch);
        if (lane == 0) {
            s_agg = agg;
            s_prefix = pre;
            if ((int64_t)len - 1 < base + SCAN_TILE) *d_count = (int)(pre + agg);
        }
    }
    __syncthreads();
    // Block-uniform values, read into scalar registers for the descriptors.
    const uint32_t pre = __builtin_amdgcn_readfirstlane(s_prefix);
    const uint32_t agg = __builtin_amdgcn_readfirstlane(s_agg);
    uint32_t dst_off[SCAN_ROWS];
#pragma unroll
    for (int k = 0; k < SCAN_ROWS; k++)
        dst_off[k] = ((mask[k] >> lane) & 1ull) ? s_wave[k][wave] + lane_rank(mask[k]) : OUT_OF_RANGE;
    if (out_index) {
        const __amdgpu_buffer_rsrc_t rd = window(out_index + pre, agg * 4);
#pragma unroll
        for (int k = 0; k < SCAN_ROWS; k++)
            BufIO<4>::st((unsigned)(base + k * SCAN_THREADS + tid), rd,
                         dst_off[k] == OUT_OF_RANGE ? OUT_OF_RANGE : dst_off[k] * 4);
    }
    for (int j = 0; j < P.n; j++) {
        const int b = P.bytes[j];
        if (b == 16)
            move_stream<16>(P.src[j], P.dst[j], base, live, pre, agg, dst_off, tid);
        else if (b == 8)
            move_stream<8>(P.src[j], P.dst[j], base, live, pre, agg, dst_off, tid);
        else
            move_stream<4>(P.src[j], P.dst[j], base, live, pre, agg, dst_off, tid);
    }
}

// CCompact.execute over the reference's records (Compact.cl:21-40 with the
// init/transfer kernels, Initialize.cl:50-79): stable compaction of 96-byte
// Intersection records (Util.cl:141-151: p, n, ng, d float4; uv float2;
// mat, id, hit int; 12 bytes padding) flagged by hit != 0, carrying the
// pixel queue along.  Tiles are 1024 records (4 rows) so each lane keeps
// its records in registers between the flag test and the move: a record is
// read from HBM once.  Survivors go to the caller's temp buffer with their
// pixel id in padding word 21; k_isects_finish writes them back in place
// and fills the tail with InitIsect records (Initialize.cl:13-23).
constexpr int ISECT_ROWS = 4;
constexpr int ISECT_TILE = SCAN_THREADS * ISECT_ROWS;
constexpr int ISECT_F4 = 6; // float4s per Intersection

__global__ __launch_bounds__(SCAN_THREADS) void k_compact_isects(const float4* __restrict__ isects,
                                                                 const int* __restrict__ pixels,
                                                                 const int* __restrict__ d_len, int capacity,
                                                                 float4* __restrict__ tmp,
                                                                 int* __restrict__ d_count, ScanWork W)
{
    __shared__ int s_tile;
    __shared__ uint32_t s_wave[ISECT_ROWS][SCAN_WAVES];
    __shared__ uint32_t s_prefix, s_agg;
    const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t epoch = W.ticket[2];
    int len = *d_len;
    if (len > capacity) len = capacity;
    if ((int64_t)blockIdx.x * ISECT_TILE >= len) return;
    if (tid == 0) s_tile = (int)atomicAdd(&W.ticket[0], 1u);
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    const int64_t base = (int64_t)tile * ISECT_TILE;
    float4 r[ISECT_ROWS][ISECT_F4];
    uint64_t mask[ISECT_ROWS];
#pragma unroll
    for (int k = 0; k < ISECT_ROWS; k++) {
        const int64_t e = base + k * SCAN_THREADS + tid;
        bool f = false;
        if (e < len) {
            const float4* src = isects + ISECT_F4 * e;
#pragma unroll
            for (int q = 0; q < ISECT_F4; q++) r[k][q] = src[q];
            f = __float_as_int(r[k][5].x) != 0; // isect->hit
        }
        mask[k] = __ballot(f);
        if (lane == 0) s_wave[k][wave] = (uint32_t)__popcll(mask[k]);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int k = 0; k < ISECT_ROWS; k++)
            for (int w = 0; w < SCAN_WAVES; w++) {
                uint32_t t = s_wave[k][w];
                s_wave[k][w] = run;
                run += t;
            }
        s_prefix = run;
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t agg = s_prefix;
        uint32_t pre = look_back(W, tile, agg, epoch);
        if (lane == 0) {
            s_agg = agg;
            s_prefix = pre;
            if ((int64_t)len - 1 < base + ISECT_TILE) *d_count = (int)(pre + agg);
        }
    }
    __syncthreads();
    const uint32_t pre = s_prefix;
#pragma unroll
    for (int k = 0; k < ISECT_ROWS; k++) {
        if ((mask[k] >> lane) & 1ull) {
            const int64_t e = base + k * SCAN_THREADS + tid;
            const uint32_t dst = pre + s_wave[k][wave] + lane_rank(mask[k]);
            float4* d = tmp + ISECT_F4 * (size_t)dst;
            float4 w5 = r[k][5];
            w5.y = __int_as_float(pixels[e]);
#pragma unroll
            for (int q = 0; q < 5; q++) d[q] = r[k][q];
            d[5] = w5;
        }
    }
}

// One lane per float4 of the first *d_len records: survivors back from the
// temp buffer (pixel id out of the padding, padding cleared), the rest
// InitIsect (p = n = ng = d = 0, uv = 0, mat = id = -1, hit = 0) with pixel 0
// (InitIntData, Initialize.cl:72-79).
__global__ void k_isects_finish(float4* __restrict__ isects, int* __restrict__ pixels,
                                const float4* __restrict__ tmp, const int* __restrict__ d_len, int capacity,
                                const int* __restrict__ d_count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int len = *d_len;
    if (len > capacity) len = capacity;
    if (i >= (int64_t)len * ISECT_F4) return;
    const int64_t j = i / ISECT_F4;
    const int q = (int)(i - j * ISECT_F4);
    float4 v;
    if (j < *d_count) {
        v = tmp[i];
        if (q == 5) {
            pixels[j] = __float_as_int(v.y);
            v = make_float4(v.x, 0.0f, 0.0f, 0.0f);
        }
    } else {
        v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (q == 4) v = make_float4(0.0f, 0.0f, __int_as_float(-1), __int_as_float(-1));
        if (q == 5) pixels[j] = 0;
    }
    isects[i] = v;
}

int32_t scan_workspace_tiles(mt_context* ctx, int64_t ntiles, ScanWork& W)
{
    size_t bytes = 256 + sizeof(uint64_t) * (size_t)(ntiles > 0 ? ntiles : 1);
    void* p = nullptr;
    int32_t rc = mt_ctx_scratch(ctx, bytes, &p);
    if (rc) return rc;
    W.ticket = reinterpret_cast<uint32_t*>(p);
    W.status = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(p) + 256);
    return MT_OK;
}

int32_t scan_workspace(mt_context* ctx, int capacity, ScanWork& W)
{
    int64_t ntiles = ((int64_t)capacity + SCAN_TILE - 1) / SCAN_TILE;
    size_t bytes = 256 + sizeof(uint64_t) * (size_t)(ntiles > 0 ? ntiles : 1);
    void* p = nullptr;
    int32_t rc = mt_ctx_scratch(ctx, bytes, &p);
    if (rc) return rc;
    W.ticket = reinterpret_cast<uint32_t*>(p);
    W.status = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(p) + 256);
    return MT_OK;
}

} // namespace

// Grows the context's look-back workspace for scans of up to `capacity`
// elements now, so later scans enqueue no allocation (stream capture).
int32_t mt_scan_reserve(mt_context* ctx, int32_t capacity)
{
    ScanWork W;
    return scan_workspace(ctx, capacity, W);
}

extern "C" int32_t mt_scan_excl_i32(mt_context* ctx, const int32_t* in, int32_t* out,
                                    const int32_t* d_len, int32_t capacity, int32_t* d_total)
{
    MT_ARG_CHECK(ctx && d_len, "mt_scan_excl_i32: NULL argument");
    MT_ARG_CHECK(capacity >= 0, "mt_scan_excl_i32: negative capacity");
    MT_ARG_CHECK(capacity == 0 || (in && out), "mt_scan_excl_i32: NULL in/out");
    ScanWork W;
    int32_t rc = scan_workspace(ctx, capacity, W);
    if (rc) return rc;
    k_scan_reset<<<1, 64, 0, ctx->stream>>>(W, d_len, capacity, d_total);
    MT_LAUNCH_CHECK();
    if (capacity == 0) return MT_OK;
    k_scan<<<mt_blocks(capacity, SCAN_TILE), SCAN_THREADS, 0, ctx->stream>>>(in, out, d_len, capacity,
                                                                          d_total, W);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

namespace {

// Argument checks and payload table shared by mt_compact and
// mt_compact_noalloc (`who` names the entry point in errors).
int32_t compact_args(const char* who, mt_context* ctx, const int32_t* flags, int32_t flag_mode,
                     const int32_t* d_len, int32_t capacity, int32_t* d_count, const mt_payload* payloads,
                     int32_t npayload, PayloadSet& P)
{
    MT_ARG_CHECK(ctx && d_len && d_count, "%s: NULL argument", who);
    MT_ARG_CHECK(capacity == 0 || flags, "%s: NULL flags", who);
    MT_ARG_CHECK(capacity >= 0, "%s: negative capacity", who);
    MT_ARG_CHECK(flag_mode == 0 || flag_mode == 1, "%s: flag_mode must be 0 or 1", who);
    MT_ARG_CHECK(npayload >= 0 && npayload <= 8, "%s: npayload %d not in [0,8]", who, npayload);
    MT_ARG_CHECK(npayload == 0 || payloads, "%s: payloads is NULL", who);
    P.n = npayload;
    for (int j = 0; j < npayload; j++) {
        MT_ARG_CHECK(capacity == 0 || (payloads[j].src && payloads[j].dst), "%s: payload %d has a NULL pointer",
                     who, j);
        int b = payloads[j].elem_bytes;
        MT_ARG_CHECK(b == 4 || b == 8 || b == 16, "%s: payload %d elem_bytes %d not 4/8/16", who, j, b);
        P.src[j] = payloads[j].src;
        P.dst[j] = payloads[j].dst;
        P.bytes[j] = b;
    }
    return MT_OK;
}

// The two launches of a compaction on look-back workspace W.
int32_t compact_launch(mt_context* ctx, const int32_t* flags, int32_t flag_mode, const int32_t* d_len,
                       int32_t capacity, int32_t* out_index, int32_t* d_count, const PayloadSet& P, ScanWork W)
{
    k_scan_reset<<<1, 64, 0, ctx->stream>>>(W, d_len, capacity, d_count);
    MT_LAUNCH_CHECK();
    if (capacity == 0) return MT_OK;
    k_compact<<<mt_blocks(capacity, SCAN_TILE), SCAN_THREADS, 0, ctx->stream>>>(
        flags, flag_mode, d_len, capacity, out_index, d_count, P, W);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

// Workspace of mt_compact_noalloc: 256 B of ticket / error flag / epoch
// words, then one 8-byte look-back status word per 8192-element tile, for
// capacities up to MT_NOALLOC_MAX_ELEMS (256 KiB in all).  Each device
// that loads the library holds its own copy until the process ends.  The
// array is zero when the code object loads: status words of epoch 0 are
// never current, the ticket is reset by k_scan_reset at every call, and the
// error flag is read and cleared by mt_scan_check.
constexpr int64_t NOALLOC_TILES = ((int64_t)MT_NOALLOC_MAX_ELEMS + SCAN_TILE - 1) / SCAN_TILE;
constexpr size_t NOALLOC_BYTES = 256 + sizeof(uint64_t) * (size_t)NOALLOC_TILES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// Device address of `scratch` on each device, resolved once per device by
// mt_noalloc_setup (host code never passes the array's own name, which is a
// host-side stand-in, to a kernel or copy).
constexpr int NOALLOC_MAX_DEVICES = 64;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES];
std::mutex g_noalloc_mu;

ScanWork noalloc_work(void* p)
{
    ScanWork W;
    W.ticket = reinterpret_cast<uint32_t*>(p);
    W.status = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(p) + 256);
    return W;
}

} // namespace

// Resolves this device's address of the no-alloc workspace; called by
// mt_ctx_create with `device` current.  Idempotent per device.
int32_t mt_noalloc_setup(int device)
{
    MT_ARG_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, "device %d beyond the %d supported", device,
                 NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return MT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        mt_set_error("hipGetSymbolAddress(no-alloc workspace) failed on device %d: %s", device,
                     hipGetErrorString(e));
        return MT_ERR_HIP;
    }
    g_noalloc_ptr[device] = p;
    return MT_OK;
}

// This device's no-alloc workspace, or nullptr before mt_noalloc_setup.
void* mt_noalloc_workspace(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

extern "C" int32_t mt_compact(mt_context* ctx, const int32_t* flags, int32_t flag_mode,
                              const int32_t* d_len, int32_t capacity, int32_t* out_index,
                              int32_t* d_count, const mt_payload* payloads, int32_t npayload)
{
    PayloadSet P;
    int32_t rc = compact_args("mt_compact", ctx, flags, flag_mode, d_len, capacity, d_count, payloads, npayload, P);
    if (rc) return rc;
    ScanWork W;
    rc = scan_workspace(ctx, capacity, W);
    if (rc) return rc;
    return compact_launch(ctx, flags, flag_mode, d_len, capacity, out_index, d_count, P, W);
}

extern "C" int32_t mt_compact_noalloc(mt_context* ctx, const int32_t* flags, int32_t flag_mode,
                                      const int32_t* d_len, int32_t capacity, int32_t* out_index,
                                      int32_t* d_count, const mt_payload* payloads, int32_t npayload)
{
    PayloadSet P;
    int32_t rc = compact_args("mt_compact_noalloc", ctx, flags, flag_mode, d_len, capacity, d_count, payloads,
                              npayload, P);
    if (rc) return rc;
    MT_ARG_CHECK(capacity <= MT_NOALLOC_MAX_ELEMS, "mt_compact_noalloc: capacity %d exceeds the maximum %d",
                 capacity, MT_NOALLOC_MAX_ELEMS);
    void* p = mt_noalloc_workspace(ctx->device);
    MT_ARG_CHECK(p != nullptr, "mt_compact_noalloc: no workspace on device %d (context not set up)", ctx->device);
    return compact_launch(ctx, flags, flag_mode, d_len, capacity, out_index, d_count, P, noalloc_work(p));
}

extern "C" int32_t mt_compact_isects(mt_context* ctx, void* isects, int32_t* pixels, void* temp_isects,
                                     const int32_t* d_len, int32_t capacity, int32_t* d_count)
{
    MT_ARG_CHECK(ctx && d_len && d_count, "mt_compact_isects: NULL argument");
    MT_ARG_CHECK(capacity >= 0, "mt_compact_isects: negative capacity");
    MT_ARG_CHECK(capacity == 0 || (isects && pixels && temp_isects), "mt_compact_isects: NULL buffer");
    {
        // the three buffers are read and written by different passes: no two may share a byte
        const uintptr_t rb = (uintptr_t)capacity * (ISECT_F4 * 16), pb = (uintptr_t)capacity * 4;
        const uintptr_t a = reinterpret_cast<uintptr_t>(isects), t = reinterpret_cast<uintptr_t>(temp_isects),
                        q = reinterpret_cast<uintptr_t>(pixels);
        auto apart = [](uintptr_t x, uintptr_t xn, uintptr_t y, uintptr_t yn) { return x + xn <= y || y + yn <= x; };
        MT_ARG_CHECK(capacity == 0 || (apart(a, rb, t, rb) && apart(q, pb, t, rb) && apart(q, pb, a, rb)),
                     "mt_compact_isects: isects, pixels and temp_isects overlap");
    }
    MT_ARG_CHECK(((reinterpret_cast<uintptr_t>(isects) | reinterpret_cast<uintptr_t>(temp_isects)) & 15) == 0,
                 "mt_compact_isects: records must be 16-byte aligned");
    ScanWork W;
    int32_t rc = scan_workspace_tiles(ctx, ((int64_t)capacity + ISECT_TILE - 1) / ISECT_TILE, W);
    if (rc) return rc;
    k_scan_reset<<<1, 64, 0, ctx->stream>>>(W, d_len, capacity, d_count);
    MT_LAUNCH_CHECK();
    if (capacity == 0) return MT_OK;
    k_compact_isects<<<mt_blocks(capacity, ISECT_TILE), SCAN_THREADS, 0, ctx->stream>>>(
        reinterpret_cast<const float4*>(isects), pixels, d_len, capacity, reinterpret_cast<float4*>(temp_isects),
        d_count, W);
    MT_LAUNCH_CHECK();
    k_isects_finish<<<mt_blocks((int64_t)capacity * ISECT_F4, 256), 256, 0, ctx->stream>>>(
        reinterpret_cast<float4*>(isects), pixels, reinterpret_cast<const float4*>(temp_isects), d_len, capacity,
        d_count);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

// Reports (and clears) the look-back spin-limit flag of the context.  A set
// flag means a scan/compaction gave up waiting and its output is invalid.
extern "C" int32_t mt_scan_check(mt_context* ctx)
{
    MT_ARG_CHECK(ctx != nullptr, "mt_scan_check: ctx is NULL");
    // the context's workspace, then the device's no-alloc workspace
    void* spaces[2] = {ctx->scratch, mt_noalloc_workspace(ctx->device)};
    for (void* sp : spaces) {
        if (!sp) continue;
        uint32_t flag = 0;
        uint32_t* tk = reinterpret_cast<uint32_t*>(sp);
        MT_HIP_CHECK(hipMemcpyAsync(&flag, tk + 1, sizeof(flag), hipMemcpyDeviceToHost, ctx->stream));
        MT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (flag) {
            int32_t rc = mt_buf_fill_i32(ctx, reinterpret_cast<int32_t*>(tk + 1), 0, 1);
            if (rc) return rc;
            mt_set_error("scan/compaction look-back exceeded its spin limit; results invalid");
            return MT_ERR_STATE;
        }
    }
    return MT_OK;
}

// identifyMeshLights + prefix sum + prepareLightInfo (LightConfiguration.cl:1-55).
__global__ void k_emitter_flags(const int* __restrict__ faces, int ntris,
                                const float* __restrict__ materials, int* __restrict__ pred)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntris) return;
    const float* m = materials + 36 * (size_t)(faces[10 * i + 9] & 0xFFFF);
    float s = m[32]; // emission_param.x (isSurfaceEmitter, Material.cl:27-31)
    float r = m[28] * s, g = m[29] * s, b = m[30] * s;
    float lum = (0.2989f * r + 0.5866f * g) + 0.1145f * b;
    pred[i] = lum != 0.0f ? 1 : 0;
}

extern "C" int32_t mt_mesh_lights(mt_context* ctx, const mt_scene* scene, const float* materials,
                                  int32_t* light_faces, int32_t* d_count)
{
    MT_ARG_CHECK(ctx && scene && materials && light_faces && d_count, "mt_mesh_lights: NULL argument");
    int n = scene->desc.num_tris;
    int* tmp = nullptr;
    MT_HIP_CHECK(hipMalloc(&tmp, sizeof(int) * ((size_t)n + 1)));
    k_emitter_flags<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(scene->desc.faces, n, materials, tmp);
    int32_t rc = mt_buf_fill_i32(ctx, tmp + n, n, 1);
    if (!rc) rc = mt_compact(ctx, tmp, 0, tmp + n, n, light_faces, d_count, nullptr, 0);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipFree(tmp);
    else (void)hipFree(tmp);
    if (rc) return rc;
    if (e != hipSuccess) {
        mt_set_error("mt_mesh_lights: %s", hipGetErrorString(e));
        return MT_ERR_HIP;
    }
    return MT_OK;
}
