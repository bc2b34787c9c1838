// Image pipeline after the bounce loop: CImage.processImage
// (src/cl/algorithms/CImage.java:124-135) = averageAccum, the log-luminance
// totals, updateFrameImage (src/cl/kernel/Image.cl:29-118, colour helpers in
// RGBSpace.cl:23-106, getIntARGB Util.cl:469-480); plus PrefixSumFloat
// (src/cl/algorithms/PrefixSumFloat.java:95-102, kernels Util.cl:982-1042)
// as a correct, deterministic float scan.
//
// Totals: the reference sums the per-pixel log luminance with its float
// prefix scan, whose global pass converts the partial sums to int
// (`int value = GET_FLOAT(...)`, Util.cl:1035) and reads aux[-1] for lane 0;
// the totals here are exact-order float sums: each 2048-element tile is
// summed as 256 lanes x 8 rows (rows in order per lane, then a shuffle-down
// tree within each wave, then the four waves in order), and the tile sums
// likewise by one work-group.  The order is fixed, so the result does not
// depend on scheduling; the oracle restates the same order.
//
// Deviations: averageAccum clears loglw/loglwcount for every launched id
// (past the image end too, Image.cl:41-42); here only pixels are written.
// xyz() returns an uninitialised colour when X+Y+Z < 1e-6 (RGBSpace.cl:70-71);
// here it returns 0.
#include "mt_internal.h"
#include "mt_devmath.h"

namespace {

constexpr int IT = 256;
constexpr int IROWS = 8;
constexpr int ITILE = IT * IROWS;

__device__ __forceinline__ float lum_ref(float r, float g, float b)
{ // Luminance, RGBSpace.cl:28-32
    return (0.2989f * r + 0.5866f * g) + 0.1145f * b;
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ int argb(float r, float g, float b, float a)
{ // getIntARGB, Util.cl:469-480
    int ri = (int)(clamp01(r) * 255.0f), gi = (int)(clamp01(g) * 255.0f);
    int bi = (int)(clamp01(b) * 255.0f), ai = (int)(clamp01(a) * 255.0f);
    return (ai << 24) | (ri << 16) | (gi << 8) | bi;
}

__global__ __launch_bounds__(IT) void k_image_init(float4* __restrict__ accum, int* __restrict__ out, int n)
{ // InitFloat4DataXYZ + InitIntDataRGB, Initialize.cl:1-11,91-98
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (accum) accum[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (out) out[i] = argb(0.0f, 0.0f, 0.0f, 1.0f);
}

__global__ __launch_bounds__(IT) void k_average(const float4* __restrict__ accum, float count, int n,
                                                float4* __restrict__ frame, float* __restrict__ loglw,
                                                int* __restrict__ logcnt)
{ // averageAccum, Image.cl:29-65
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = accum[i];
    float4 f = make_float4(a.x / count, a.y / count, a.z / count, 1.0f);
    frame[i] = f;
    float lw = lum_ref(f.x, f.y, f.z);
    int p = 0;
    float v = 0.0f;
    if (lw > 0.0f) {
        v = logf(0.01f + lw);
        p = 1;
    }
    loglw[i] = v;
    logcnt[i] = p;
}

// Fixed-order sum of one tile (see the header): returns the sum on thread 0.
template <typename T>
__device__ __forceinline__ T tile_sum(const T* __restrict__ x, int64_t base, int64_t n, T* s_wave)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T s = 0;
#pragma unroll
    for (int k = 0; k < IROWS; k++) {
        int64_t i = base + (int64_t)k * IT + tid;
        s = s + (i < n ? x[i] : (T)0);
    }
    for (int o = 32; o > 0; o >>= 1) s = s + __shfl_down(s, o);
    if (lane == 0) s_wave[wave] = s;
    __syncthreads();
    T t = 0;
    if (tid == 0) t = ((s_wave[0] + s_wave[1]) + s_wave[2]) + s_wave[3];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(IT) void k_tile_sums(const float* __restrict__ f, const int* __restrict__ c, int n,
                                                  float* __restrict__ fpart, int* __restrict__ cpart)
{
    __shared__ float sf[4];
    __shared__ int sc[4];
    const int64_t base = (int64_t)blockIdx.x * ITILE;
    float a = tile_sum(f, base, n, sf);
    int b = c ? tile_sum(c, base, n, sc) : 0;
    if (threadIdx.x == 0) {
        fpart[blockIdx.x] = a;
        if (c) cpart[blockIdx.x] = b;
    }
}

// One work-group: tile partials summed in the same order (lane t takes
// partials t, t+256, ... in order).
__global__ __launch_bounds__(IT) void k_final_sums(const float* __restrict__ fpart, const int* __restrict__ cpart,
                                                   int ntiles, float* __restrict__ ftot, int* __restrict__ ctot)
{
    __shared__ float sf[4];
    __shared__ int sc[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float a = 0.0f;
    int b = 0;
    for (int i = tid; i < ntiles; i += IT) {
        a = a + fpart[i];
        if (cpart) b = b + cpart[i];
    }
    for (int o = 32; o > 0; o >>= 1) {
        a = a + __shfl_down(a, o);
        b = b + __shfl_down(b, o);
    }
    if (lane == 0) {
        sf[wave] = a;
        sc[wave] = b;
    }
    __syncthreads();
    if (tid == 0) {
        *ftot = ((sf[0] + sf[1]) + sf[2]) + sf[3];
        if (ctot) *ctot = ((sc[0] + sc[1]) + sc[2]) + sc[3];
    }
}

__global__ __launch_bounds__(IT) void k_update(float4* __restrict__ frame, int* __restrict__ out, int n,
                                               const float* __restrict__ tot_log, const int* __restrict__ tot_cnt,
                                               float exposure, float gamma)
{ // updateFrameImage, Image.cl:68-118
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float avg = expf(*tot_log / (float)*tot_cnt);
    if (i >= n) return;
    float4 c = frame[i];
    float cl = lum_ref(c.x, c.y, c.z);
    if (!(cl > 0.0f)) return;
    float scaled = (exposure * cl) / avg;
    float Y = scaled / (1.0f + scaled); // toneSimpleReinhard, RGBSpace.cl:23-26
    if (isinf(Y) || isnan(Y) || Y < 0.000001f) Y = 0.0f; // check, Util.cl:381-387
    // convertRGBtoXYZ, RGBSpace.cl:34-47
    float X = (0.5893f * c.x + 0.1789f * c.y) + 0.1831f * c.z;
    float Yc = (0.2904f * c.x + 0.6051f * c.y) + 0.1045f * c.z;
    float Z = (0.0000f * c.x + 0.0684f * c.y) + 1.0202f * c.z;
    // xyz, RGBSpace.cl:65-78
    float sum = (X + Yc) + Z;
    float x = 0.0f, y = 0.0f;
    if (!(sum < 1e-6f)) {
        float s = 1.0f / sum;
        x = X * s;
        y = Yc * s;
    }
    // xyYtoXYZ, RGBSpace.cl:80-106
    float X2 = y < 1e-6f ? 0.0f : Y * (x / y);
    float Z2 = y < 1e-6f ? 0.0f : (Y * ((1.0f - x) - y)) / y;
    // convertXYZtoRGB, RGBSpace.cl:49-63
    float r = (1.967f * X2 - 0.548f * Y) - 0.297f * Z2;
    float g = (-0.955f * X2 + 1.938f * Y) - 0.027f * Z2;
    float b = (0.064f * X2 - 0.130f * Y) + 0.982f * Z2;
    float ig = 1.0f / gamma;
    r = powf(r, ig);
    g = powf(g, ig);
    b = powf(b, ig);
    frame[i] = make_float4(r, g, b, 1.0f);
    out[i] = argb(r, g, b, 1.0f);
}

// ---------------------------------------------------------------- float scan
__global__ __launch_bounds__(IT) void k_fscan_tiles(const float* __restrict__ in, const int* __restrict__ d_len,
                                                    int capacity, float* __restrict__ part)
{
    __shared__ float sf[4];
    int n = *d_len;
    n = n > capacity ? capacity : n;
    const int64_t base = (int64_t)blockIdx.x * ITILE;
    if (base >= n) {
        if (threadIdx.x == 0) part[blockIdx.x] = 0.0f;
        return;
    }
    float a = tile_sum(in, base, n, sf);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// Exclusive scan of the tile sums, in order, by one thread per 256-chunk
// then sequentially across chunks (a single work-group; tiles <= 2^20).
__global__ __launch_bounds__(IT) void k_fscan_parts(float* __restrict__ part, int ntiles, const int* __restrict__ d_len,
                                                    int capacity, float* __restrict__ d_total)
{
    __shared__ float s_chunk[IT];
    const int tid = threadIdx.x;
    const int per = (ntiles + IT - 1) / IT;
    const int b = tid * per, e = b + per < ntiles ? b + per : ntiles;
    float run = 0.0f;
    for (int i = b; i < e; i++) run = run + part[i];
    s_chunk[tid] = run;
    __syncthreads();
    if (tid == 0) { // chunk offsets in order
        float acc = 0.0f;
        for (int k = 0; k < IT; k++) {
            float v = s_chunk[k];
            s_chunk[k] = acc;
            acc = acc + v;
        }
        int n = *d_len;
        n = n > capacity ? capacity : n;
        if (d_total) *d_total = n > 0 ? acc : 0.0f;
    }
    __syncthreads();
    float acc = s_chunk[tid];
    for (int i = b; i < e; i++) {
        float v = part[i];
        part[i] = acc;
        acc = acc + v;
    }
}

// Within a tile: each lane's 8 elements are consecutive (lane-major within
// the tile), lane totals scanned across the tile in order.
__global__ __launch_bounds__(IT) void k_fscan_apply(const float* __restrict__ in, float* __restrict__ out,
                                                    const int* __restrict__ d_len, int capacity,
                                                    const float* __restrict__ part)
{
    __shared__ float s_lane[IT];
    int n = *d_len;
    n = n > capacity ? capacity : n;
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * ITILE;
    if (base >= n) return;
    const int64_t first = base + (int64_t)tid * IROWS;
    float v[IROWS];
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < IROWS; k++) {
        v[k] = first + k < n ? in[first + k] : 0.0f;
        s = s + v[k];
    }
    s_lane[tid] = s;
    __syncthreads();
    if (tid == 0) {
        float acc = part[blockIdx.x];
        for (int k = 0; k < IT; k++) {
            float t = s_lane[k];
            s_lane[k] = acc;
            acc = acc + t;
        }
    }
    __syncthreads();
    float acc = s_lane[tid];
#pragma unroll
    for (int k = 0; k < IROWS; k++) {
        if (first + k < n) out[first + k] = acc;
        acc = acc + v[k];
    }
}

} // namespace

extern "C" int32_t mt_image_init(mt_context* ctx, float* accum, int32_t* argb_out, int32_t n)
{
    MT_ARG_CHECK(ctx && n >= 0, "mt_image_init: bad arguments");
    if (n == 0) return MT_OK;
    k_image_init<<<mt_blocks(n, IT), IT, 0, ctx->stream>>>(reinterpret_cast<float4*>(accum), argb_out, n);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_image_process(mt_context* ctx, const float* accum, float frame_count, int32_t n,
                                    float exposure, float gamma, int32_t update_totals, float* frame,
                                    int32_t* argb_out, float* d_total_log, int32_t* d_total_count)
{
    MT_ARG_CHECK(ctx && accum && frame && argb_out && d_total_log && d_total_count,
                 "mt_image_process: NULL argument");
    MT_ARG_CHECK(n > 0 && n <= (1 << 30), "mt_image_process: n = %d", n);
    hipStream_t s = ctx->stream;
    const int grid = mt_blocks(n, IT);
    const int ntiles = (int)(((int64_t)n + ITILE - 1) / ITILE);
    // per-pixel log luminance and predicate, then the tile partials, in the
    // context workspace
    size_t need = (size_t)n * 8 + (size_t)ntiles * 8;
    void* w = nullptr;
    int32_t rc = mt_ctx_temp(ctx, need, &w);
    if (rc) return rc;
    char* base = reinterpret_cast<char*>(w);
    float* loglw = reinterpret_cast<float*>(base);
    int* cnt = reinterpret_cast<int*>(base + (size_t)n * 4);
    float* fpart = reinterpret_cast<float*>(base + (size_t)n * 8);
    int* cpart = reinterpret_cast<int*>(base + (size_t)n * 8 + (size_t)ntiles * 4);
    k_average<<<grid, IT, 0, s>>>(reinterpret_cast<const float4*>(accum), frame_count, n,
                                  reinterpret_cast<float4*>(frame), loglw, cnt);
    MT_LAUNCH_CHECK();
    if (update_totals) { // CImage.processImage recomputes them for the first frames only
        k_tile_sums<<<ntiles, IT, 0, s>>>(loglw, cnt, n, fpart, cpart);
        k_final_sums<<<1, IT, 0, s>>>(fpart, cpart, ntiles, d_total_log, d_total_count);
        MT_LAUNCH_CHECK();
    }
    k_update<<<grid, IT, 0, s>>>(reinterpret_cast<float4*>(frame), argb_out, n, d_total_log, d_total_count,
                                 exposure, gamma);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_scan_excl_f32(mt_context* ctx, const float* in, float* out, const int32_t* d_len,
                                    int32_t capacity, float* d_total)
{
    MT_ARG_CHECK(ctx && d_len, "mt_scan_excl_f32: NULL argument");
    MT_ARG_CHECK(capacity >= 0 && capacity <= (1 << 30), "mt_scan_excl_f32: capacity %d", capacity);
    MT_ARG_CHECK(capacity == 0 || (in && out), "mt_scan_excl_f32: NULL in/out");
    const int ntiles = capacity > 0 ? (int)(((int64_t)capacity + ITILE - 1) / ITILE) : 1;
    void* w = nullptr;
    int32_t rc = mt_ctx_temp(ctx, (size_t)ntiles * 4, &w);
    if (rc) return rc;
    float* part = reinterpret_cast<float*>(w);
    hipStream_t s = ctx->stream;
    if (capacity > 0) {
        k_fscan_tiles<<<ntiles, IT, 0, s>>>(in, d_len, capacity, part);
        MT_LAUNCH_CHECK();
    }
    k_fscan_parts<<<1, IT, 0, s>>>(part, capacity > 0 ? ntiles : 0, d_len, capacity, d_total);
    MT_LAUNCH_CHECK();
    if (capacity > 0) {
        k_fscan_apply<<<ntiles, IT, 0, s>>>(in, out, d_len, capacity, part);
        MT_LAUNCH_CHECK();
    }
    return MT_OK;
}
