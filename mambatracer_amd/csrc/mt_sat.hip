// Stage C entry points: adaptive light-grid SAT build and sampling, and the
// environment-map luminance/SAT and light sampling.
//
// Light-grid build.  The reference rebuilds the 1600x1600 grid's tile SATs
// with four launches over 2.56M work-items (initFuncAndSat, calculateFunc,
// prefixSumRowSubgrid with 16-wide work-groups, prefixSumColSubgrid with
// 32-wide work-groups and strided column reads; LightGrid.cl:347-423,
// CAdaptiveEnvironment.java:135-145).  Here one wave owns one 16x32 tile:
// it reads the tile's 32 rows (a work-group of 4 waves covers 4 adjacent
// tiles, so every row access is a contiguous 256 B), runs the 16-wide
// Kogge-Stone row scans with lane shuffles, transposes through LDS, runs the
// 32-tall Kogge-Stone column scans with shuffles, and writes the SAT once.
// Each scan step computes x[l] = x[l-o] + x[l] for l >= o, o = 1, 2, 4, ...,
// the reference's exact summation tree, so the SAT is bit-identical.
// HBM traffic: read func (or accum) 4 B + write sat 4 B (+ func 4 B when
// updating from accum) per cell.
#include "mt_sat.h"

namespace {

constexpr int LDS_STRIDE = LG_RX + 1; // padded row stride for the column pass

// mode 0: src = func (build_sat); mode 1: src = accum, also func <- accum (update)
__global__ __launch_bounds__(256) void k_lg_build(const float* __restrict__ src,
                                                  float* __restrict__ func,
                                                  float* __restrict__ sat, int mode)
{
    __shared__ float lds[4][LG_RY * LDS_STRIDE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = blockIdx.x * 4 + wave;
    if (s >= LG_CELLS) return; // LG_CELLS is a multiple of 4: whole blocks exit together
    const int x0 = (s % LG_CX) * LG_RX, y0 = (s / LG_CX) * LG_RY;
    float* L = lds[wave];
    const int c = lane & 15, r4 = lane >> 4;
#pragma unroll
    for (int j = 0; j < LG_RY / 4; j++) {
        int r = 4 * j + r4;
        size_t g = (size_t)(y0 + r) * LG_NU + x0 + c;
        float x = src[g];
        if (mode == 1) func[g] = x;
#pragma unroll
        for (int o = 1; o < LG_RX; o <<= 1) {
            float y = __shfl_up(x, o, 16);
            if (c >= o) x = y + x;
        }
        L[r * LDS_STRIDE + c] = x;
    }
    __syncthreads();
    const int rr = lane & 31, half = lane >> 5;
#pragma unroll
    for (int j = 0; j < LG_RX / 2; j++) {
        int cc = 2 * j + half;
        float x = L[rr * LDS_STRIDE + cc];
#pragma unroll
        for (int o = 1; o < LG_RY; o <<= 1) {
            float y = __shfl_up(x, o, 32);
            if (rr >= o) x = y + x;
        }
        L[rr * LDS_STRIDE + cc] = x;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LG_RY / 4; j++) {
        int r = 4 * j + r4;
        sat[(size_t)(y0 + r) * LG_NU + x0 + c] = L[r * LDS_STRIDE + c];
    }
}

__global__ void k_lg_clear(float* __restrict__ a, float* __restrict__ b, float* __restrict__ c,
                           unsigned long long* __restrict__ pend)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < LG_SIZE) {
        if (a) a[i] = 0.0f;
        b[i] = 0.0f;
        c[i] = 0.0f;
        if (pend) {
            for (int k = 0; k < LG_PEND_COPIES; k++) pend[(size_t)k * LG_SIZE + i] = 0ull;
        }
    }
}

// End of a render in the guided environment mode: the frame's fixed-point
// sums join the float accumulation, one rounding per entry, and are cleared.
__global__ void k_lg_fold(unsigned long long* __restrict__ pend, float* __restrict__ accum)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= LG_SIZE) return;
    unsigned long long q = 0ull;
    for (int k = 0; k < LG_PEND_COPIES; k++) {
        const unsigned long long v = pend[(size_t)k * LG_SIZE + i];
        if (v) {
            q += v;
            pend[(size_t)k * LG_SIZE + i] = 0ull;
        }
    }
    if (q) accum[i] = accum[i] + (float)((double)q * (1.0 / 1048576.0));
}

__global__ void k_lg_set_tile(float* __restrict__ func, const float* __restrict__ tile)
{ // setToAllFunc, CLightGrid.java:86-100: entry i of the tile at local (i % 16, i / 16)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= LG_SIZE) return;
    int s = i / (LG_RX * LG_RY), k = i % (LG_RX * LG_RY);
    func[lg_global_index(s, k % LG_RX, k / LG_RX)] = tile[k];
}

// Generic tile SAT (test sizes): row segments of rx, then column segments of ry.
__global__ void k_sat_rows(const float* __restrict__ func, float* __restrict__ sat, int64_t nseg, int rx)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nseg) return;
    float a[64];
    for (int l = 0; l < rx; l++) a[l] = func[g * rx + l];
    for (int o = 1; o < rx; o *= 2)
        for (int l = rx - 1; l >= o; l--) a[l] = a[l - o] + a[l];
    for (int l = 0; l < rx; l++) sat[g * rx + l] = a[l];
}

__global__ void k_sat_cols(float* __restrict__ sat, int64_t nseg, int nu, int nv, int ry)
{ // prefixSumColSubgrid indexing, LightGrid.cl:379-383
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nseg) return;
    int64_t i = (int64_t)ry * g;
    int64_t xi = i / nv, yi0 = i % nv;
    float a[64];
    for (int l = 0; l < ry; l++) a[l] = sat[xi + (yi0 + l) * nu];
    for (int o = 1; o < ry; o *= 2)
        for (int l = ry - 1; l >= o; l--) a[l] = a[l - o] + a[l];
    for (int l = 0; l < ry; l++) sat[xi + (yi0 + l) * nu] = a[l];
}

__global__ void k_env_tile_tables(EnvView E, float* __restrict__ ttab)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)LG_RX * LG_RY * E.tstride) return;
    const int k = (int)(i / E.tstride);
    ttab[i] = env_tile_entry(E, k, (int)(i - (int64_t)k * E.tstride));
}

// One work-group per cell writes its sampling tables.
__global__ __launch_bounds__(256) void k_lg_tables(const float* __restrict__ sat, float* __restrict__ tab)
{
    const int s = blockIdx.x;
    for (int i = threadIdx.x; i < LG_TAB; i += blockDim.x) tab[(size_t)s * LG_TAB + i] = lg_table_entry(sat, s, i);
}

__global__ void k_lg_sample(const float* __restrict__ func, const float* __restrict__ sat,
                            const float* __restrict__ tab, int n, const int* __restrict__ cell,
                            const float2* __restrict__ u, int2* __restrict__ tile_xy, float2* __restrict__ uv,
                            float* __restrict__ pdf)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = cell[i];
    if (s < 0 || s >= LG_CELLS) {
        tile_xy[i] = make_int2(-1, -1);
        uv[i] = make_float2(0, 0);
        pdf[i] = 0.0f;
        return;
    }
    float2 uu = u[i];
    int ox, oy;
    float a, b, p;
    if (tab) lg_sample_tab(func, tab, s, uu.x, uu.y, ox, oy, a, b, p);
    else lg_sample(func, sat, s, uu.x, uu.y, ox, oy, a, b, p);
    tile_xy[i] = make_int2(ox, oy);
    uv[i] = make_float2(a, b);
    pdf[i] = p;
}

__global__ void k_lg_cells(int n, const float4* __restrict__ dirs, int* __restrict__ cell,
                           int2* __restrict__ tile_xy)
{ // subgridIndexFromCamera / tileIndexXYFromCamera, LightGrid.cl:68-90
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v3 v = xyz(dirs[i]);
    cell[i] = sph_index(LG_CX, LG_CY, norm3(v));
    int t = sph_index(LG_RX, LG_RY, v);
    tile_xy[i] = make_int2(t % LG_RX, t / LG_RX);
}

// ---------------------------------------------------------------- environment map
__global__ void k_env_lum(const float4* __restrict__ rgba, int64_t n, float* __restrict__ lum,
                          float* __restrict__ sat)
{ // InitEnvironmentLum, Initialize.cl:100-115
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = rgba[i];
    float l = luminance3(c.x, c.y, c.z);
    lum[i] = l;
    sat[i] = l;
}

// calculateSAT rows (SAT.cl:114-122,306-307): one thread per row carries the
// running sum left to right; 64x64 blocks are staged through LDS so global
// accesses stay coalesced.
__global__ __launch_bounds__(64) void k_env_rows(float* __restrict__ sat, int w, int h)
{
    __shared__ float t[64][65];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * 64;
    float val = 0.0f;
    for (int x0 = 0; x0 < w; x0 += 64) {
        for (int r = 0; r < 64; r++) {
            int y = y0 + r, x = x0 + tid;
            t[r][tid] = (y < h && x < w) ? sat[(size_t)y * w + x] : 0.0f;
        }
        __syncthreads();
        int lim = w - x0 < 64 ? w - x0 : 64;
        for (int k = 0; k < lim; k++) {
            val = val + t[tid][k];
            t[tid][k] = val;
        }
        __syncthreads();
        for (int r = 0; r < 64; r++) {
            int y = y0 + r, x = x0 + tid;
            if (y < h && x < w) sat[(size_t)y * w + x] = t[r][tid];
        }
        __syncthreads();
    }
}

// calculateSAT columns (SAT.cl:125-133,308-309): one thread per column; lanes
// of a wave walk adjacent columns, so each step is a coalesced row access.
__global__ __launch_bounds__(256) void k_env_cols(float* __restrict__ sat, int w, int h)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    float val = 0.0f;
    int y = 0;
    for (; y + 8 <= h; y += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = sat[(size_t)(y + k) * w + x];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            val = val + v[k];
            sat[(size_t)(y + k) * w + x] = val;
        }
    }
    for (; y < h; y++) {
        val = val + sat[(size_t)y * w + x];
        sat[(size_t)y * w + x] = val;
    }
}

// Exact full-map sampling tables: the same expressions the sampler evaluates
// from SAT corners, stored once (marg/fint/pmarg: h+1 entries, cond: h rows
// of w+1).
__global__ void k_env_tables(const float* __restrict__ sat, int w, int h, float* __restrict__ marg,
                             float* __restrict__ cond, float* __restrict__ fint, float* __restrict__ pmarg)
{
    SatRegion R{w, h, 0, 0, w, h};
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ncond = (int64_t)h * (w + 1);
    if (i < ncond) {
        int y = (int)(i / (w + 1)), x = (int)(i % (w + 1));
        cond[i] = r_cond(R, sat, x, y);
    } else if (i < ncond + h + 1) {
        int y = (int)(i - ncond);
        marg[y] = r_marg(R, sat, y);
        fint[y] = r_func_int(R, sat, y);
        pmarg[y] = r_pdf_marg(R, sat, y);
    }
}

// Search guides over the tables above: gmarg[k] = cdf_search(marg, 0, h,
// k / ENV_GUIDE_M) for k in [0, ENV_GUIDE_M + 1]; row y of gcond likewise
// over cond row y (w entries) with ENV_GUIDE_C intervals.
__global__ void k_env_guides(const float* __restrict__ marg, const float* __restrict__ cond, int w, int h,
                             int* __restrict__ gmarg, int* __restrict__ gcond)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nm = ENV_GUIDE_M + 2, nc = (int64_t)h * (ENV_GUIDE_C + 2);
    if (i < nm) {
        gmarg[i] = cdf_search(marg, 0, h, (float)i / (float)ENV_GUIDE_M);
    } else if (i < nm + nc) {
        const int64_t j = i - nm;
        const int y = (int)(j / (ENV_GUIDE_C + 2)), k = (int)(j % (ENV_GUIDE_C + 2));
        gcond[j] = cdf_search(cond + (size_t)y * (w + 1), 0, w, (float)k / (float)ENV_GUIDE_C);
    }
}

__global__ void k_region_sample(const float* __restrict__ func, const float* __restrict__ sat, SatRegion R,
                                int n, const float2* __restrict__ u, float2* __restrict__ uv,
                                int2* __restrict__ off, float* __restrict__ pdf)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 q = u[i];
    float a, b, p;
    int ox, oy;
    r_sample(R, sat, func, q.x, q.y, a, b, ox, oy, p);
    uv[i] = make_float2(a, b);
    off[i] = make_int2(ox, oy);
    pdf[i] = p;
}

__global__ void k_env_illuminate(EnvView E, int n, const float2* __restrict__ u, float4* __restrict__ dir,
                                 float4* __restrict__ contrib, float* __restrict__ pdfw,
                                 int* __restrict__ texel)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 q = u[i];
    v3 d;
    float pw;
    int tx;
    float4 c = env_illuminate(E, q.x, q.y, d, pw, tx);
    dir[i] = make_float4(d.x, d.y, d.z, 0.0f);
    contrib[i] = c;
    pdfw[i] = pw;
    texel[i] = tx;
}

__global__ void k_env_radiance(EnvView E, int n, const float4* __restrict__ dir, float4* __restrict__ contrib,
                               float* __restrict__ pdfw, int* __restrict__ texel)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pw;
    int tx;
    float4 c = env_radiance(E, xyz(dir[i]), pw, tx);
    contrib[i] = c;
    pdfw[i] = pw;
    texel[i] = tx;
}

__global__ void k_env_illuminate_adaptive(EnvView E, const float* __restrict__ lg_func,
                                          const float* __restrict__ lg_sat_arr, const float* __restrict__ lg_tab,
                                          float cx, float cy, float cz, int n,
                                          const float4* __restrict__ hit, const float4* __restrict__ u,
                                          float4* __restrict__ dir, float4* __restrict__ contrib,
                                          float* __restrict__ pdfw, float* __restrict__ lum, int* __restrict__ gidx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v3 d;
    float pw, l;
    int g;
    float tp;
    const float4 c =
        env_illuminate_adaptive(E, lg_func, lg_sat_arr, lg_tab, mk3(cx, cy, cz), xyz(hit[i]), u[i], d, pw, l, g, tp);
    dir[i] = make_float4(d.x, d.y, d.z, 0.0f);
    contrib[i] = c;
    pdfw[i] = pw;
    if (lum) lum[i] = l;
    if (gidx) gidx[i] = g;
}

// accumLightGrid (Light.cl:35-38) adds each luminance with a float atomic, as
// the reference's compare-and-swap loop (Util.cl:783-795) does: the sum per
// entry is exact up to the order of the additions.
__global__ void k_env_radiance_adaptive(EnvView E, float* accum, float cx, float cy, float cz, int n,
                                        const float4* __restrict__ dir, const float4* __restrict__ hit,
                                        float4* __restrict__ contrib, float* __restrict__ pdfa,
                                        float* __restrict__ lum, int* __restrict__ gidx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pa, l;
    int g, s, t;
    const float4 c = env_radiance_adaptive(E, mk3(cx, cy, cz), xyz(dir[i]), xyz(hit[i]), pa, l, g, s, t);
    if (accum) atomicAdd(accum + g, l);
    if (contrib) contrib[i] = c;
    if (pdfa) pdfa[i] = pa;
    if (lum) lum[i] = l;
    if (gidx) gidx[i] = g;
}

} // namespace

extern "C" int32_t mt_lightgrid_create(mt_context* ctx, mt_lightgrid** out)
{
    MT_ARG_CHECK(ctx && out, "mt_lightgrid_create: NULL argument");
    *out = nullptr;
    mt_lightgrid* g = new mt_lightgrid();
    g->ctx = ctx;
    float* p = nullptr;
    if (hipMalloc(&p, sizeof(float) * 3 * (size_t)LG_SIZE) != hipSuccess) {
        delete g;
        mt_set_error("mt_lightgrid_create: device allocation failed");
        return MT_ERR_ALLOC;
    }
    g->accum = p;
    g->func = p + LG_SIZE;
    g->sat = p + 2 * (size_t)LG_SIZE;
    if (hipMalloc(&g->tab, sizeof(float) * LG_TAB * (size_t)LG_CELLS) != hipSuccess) {
        (void)hipFree(p);
        delete g;
        mt_set_error("mt_lightgrid_create: device allocation failed");
        return MT_ERR_ALLOC;
    }
    g->tab_valid = false;
    if (hipMalloc(&g->pend, sizeof(unsigned long long) * (size_t)LG_PEND_COPIES * LG_SIZE) != hipSuccess) {
        (void)hipFree(p);
        (void)hipFree(g->tab);
        delete g;
        mt_set_error("mt_lightgrid_create: device allocation failed");
        return MT_ERR_ALLOC;
    }
    *out = g;
    return mt_lightgrid_clear(ctx, g);
}

extern "C" int32_t mt_lightgrid_destroy(mt_lightgrid* g)
{
    if (!g) return MT_OK;
    MtFirstError fe;
    fe(hipStreamSynchronize(g->ctx->stream));
    fe(hipFree(g->accum));
    fe(hipFree(g->tab));
    fe(hipFree(g->pend));
    delete g;
    return fe.status("mt_lightgrid_destroy");
}

extern "C" int32_t mt_lightgrid_arrays(mt_lightgrid* g, float** accum, float** func, float** sat)
{
    MT_ARG_CHECK(g != nullptr, "mt_lightgrid_arrays: NULL grid");
    // the caller may now write func/sat directly: sample from the SAT itself
    // until the next SAT build refreshes the tables
    if (func || sat) g->tab_valid = false;
    if (accum) *accum = g->accum;
    if (func) *func = g->func;
    if (sat) *sat = g->sat;
    return MT_OK;
}

// Refreshes the sampling tables after the SAT changed (stream-ordered).
static int32_t lg_refresh_tables(mt_context* ctx, mt_lightgrid* g)
{
    k_lg_tables<<<LG_CELLS, 256, 0, ctx->stream>>>(g->sat, g->tab);
    MT_LAUNCH_CHECK();
    g->tab_valid = true;
    return MT_OK;
}

extern "C" int32_t mt_lightgrid_clear(mt_context* ctx, mt_lightgrid* g)
{
    MT_ARG_CHECK(ctx && g, "mt_lightgrid_clear: NULL argument");
    k_lg_clear<<<mt_blocks(LG_SIZE, 256), 256, 0, ctx->stream>>>(g->accum, g->func, g->sat, g->pend);
    MT_LAUNCH_CHECK();
    return lg_refresh_tables(ctx, g);
}

extern "C" int32_t mt_lightgrid_set_tile_func(mt_context* ctx, mt_lightgrid* g, const float* tile)
{
    MT_ARG_CHECK(ctx && g && tile, "mt_lightgrid_set_tile_func: NULL argument");
    float* dtile = nullptr;
    MT_HIP_CHECK(hipMalloc(&dtile, sizeof(float) * LG_RX * LG_RY));
    hipError_t e = hipMemcpyAsync(dtile, tile, sizeof(float) * LG_RX * LG_RY, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) {
        k_lg_set_tile<<<mt_blocks(LG_SIZE, 256), 256, 0, ctx->stream>>>(g->func, dtile);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        k_lg_build<<<LG_CELLS / 4, 256, 0, ctx->stream>>>(g->func, g->func, g->sat, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        k_lg_tables<<<LG_CELLS, 256, 0, ctx->stream>>>(g->sat, g->tab);
        e = hipGetLastError();
        g->tab_valid = e == hipSuccess;
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    if (e2 == hipSuccess) e2 = hipFree(dtile);
    else (void)hipFree(dtile);
    if (e != hipSuccess || e2 != hipSuccess) {
        mt_set_error("mt_lightgrid_set_tile_func: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        return MT_ERR_HIP;
    }
    return MT_OK;
}

extern "C" int32_t mt_lightgrid_update(mt_context* ctx, mt_lightgrid* g)
{
    MT_ARG_CHECK(ctx && g, "mt_lightgrid_update: NULL argument");
    k_lg_build<<<LG_CELLS / 4, 256, 0, ctx->stream>>>(g->accum, g->func, g->sat, 1);
    MT_LAUNCH_CHECK();
    return lg_refresh_tables(ctx, g);
}

extern "C" int32_t mt_lightgrid_build_sat(mt_context* ctx, mt_lightgrid* g)
{
    MT_ARG_CHECK(ctx && g, "mt_lightgrid_build_sat: NULL argument");
    k_lg_build<<<LG_CELLS / 4, 256, 0, ctx->stream>>>(g->func, g->func, g->sat, 0);
    MT_LAUNCH_CHECK();
    return lg_refresh_tables(ctx, g);
}

extern "C" int32_t mt_lightgrid_sample(mt_context* ctx, const mt_lightgrid* g, int32_t n,
                                       const int32_t* cell, const float* u, int32_t* tile_xy,
                                       float* uv, float* pdf)
{
    MT_ARG_CHECK(ctx && g && cell && u && tile_xy && uv && pdf, "mt_lightgrid_sample: NULL argument");
    MT_ARG_CHECK(n >= 0, "mt_lightgrid_sample: negative n");
    if (n == 0) return MT_OK;
    k_lg_sample<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(g->func, g->sat, g->tab_valid ? g->tab : nullptr, n, cell,
                                                            reinterpret_cast<const float2*>(u),
                                                            reinterpret_cast<int2*>(tile_xy),
                                                            reinterpret_cast<float2*>(uv), pdf);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_lightgrid_cells(mt_context* ctx, int32_t n, const float* dirs, int32_t* cell,
                                      int32_t* tile_xy)
{
    MT_ARG_CHECK(ctx && dirs && cell && tile_xy, "mt_lightgrid_cells: NULL argument");
    MT_ARG_CHECK(n >= 0, "mt_lightgrid_cells: negative n");
    if (n == 0) return MT_OK;
    k_lg_cells<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(n, reinterpret_cast<const float4*>(dirs), cell,
                                                           reinterpret_cast<int2*>(tile_xy));
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_sat_tiles(mt_context* ctx, const float* func, float* sat, int32_t nu, int32_t nv,
                                int32_t rx, int32_t ry)
{
    MT_ARG_CHECK(ctx && func && sat, "mt_sat_tiles: NULL argument");
    MT_ARG_CHECK(nu > 0 && nv > 0 && rx > 0 && ry > 0 && rx <= 64 && ry <= 64,
                 "mt_sat_tiles: bad sizes %d x %d tiles %d x %d", nu, nv, rx, ry);
    int64_t total = (int64_t)nu * nv;
    MT_ARG_CHECK(total % rx == 0 && total % ry == 0 && nv % ry == 0,
                 "mt_sat_tiles: tiles must divide the grid");
    k_sat_rows<<<mt_blocks(total / rx, 128), 128, 0, ctx->stream>>>(func, sat, total / rx, rx);
    MT_LAUNCH_CHECK();
    k_sat_cols<<<mt_blocks(total / ry, 128), 128, 0, ctx->stream>>>(sat, total / ry, nu, nv, ry);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_env_lum_sat(mt_context* ctx, const float* rgba, int32_t w, int32_t h, float* lum,
                                  float* sat)
{
    MT_ARG_CHECK(ctx && rgba && lum && sat, "mt_env_lum_sat: NULL argument");
    MT_ARG_CHECK(w > 0 && h > 0 && (int64_t)w * h < (1ll << 31), "mt_env_lum_sat: bad size %d x %d", w, h);
    int64_t n = (int64_t)w * h;
    k_env_lum<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(reinterpret_cast<const float4*>(rgba), n, lum, sat);
    MT_LAUNCH_CHECK();
    k_env_rows<<<mt_blocks(h, 64), 64, 0, ctx->stream>>>(sat, w, h);
    MT_LAUNCH_CHECK();
    k_env_cols<<<mt_blocks(w, 256), 256, 0, ctx->stream>>>(sat, w, h);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_env_create(mt_context* ctx, const float* rgba, const float* lum, const float* sat,
                                 int32_t w, int32_t h, mt_envmap** out)
{
    MT_ARG_CHECK(ctx && rgba && lum && sat && out, "mt_env_create: NULL argument");
    MT_ARG_CHECK(w > 0 && h > 0, "mt_env_create: bad size");
    mt_envmap* e = new mt_envmap();
    e->ctx = ctx;
    e->rgba = rgba;
    e->lum = lum;
    e->sat = sat;
    e->w = w;
    e->h = h;
    int64_t ncond = (int64_t)h * (w + 1);
    const int64_t ngm = ENV_GUIDE_M + 2, ngc = (int64_t)h * (ENV_GUIDE_C + 2);
    size_t nf = (size_t)ncond + 3 * (size_t)(h + 1) + (size_t)(ngm + ngc);
    if (hipMalloc(&e->tab, sizeof(float) * nf) != hipSuccess) {
        delete e;
        mt_set_error("mt_env_create: cannot allocate sampling tables");
        return MT_ERR_ALLOC;
    }
    e->cond = e->tab;
    e->marg = e->tab + ncond;
    e->fint = e->marg + (h + 1);
    e->pmarg = e->fint + (h + 1);
    e->gmarg = reinterpret_cast<int*>(e->pmarg + (h + 1));
    e->gcond = e->gmarg + ngm;
    k_env_tables<<<mt_blocks(ncond + h + 1, 256), 256, 0, ctx->stream>>>(sat, w, h, e->marg, e->cond, e->fint,
                                                                         e->pmarg);
    k_env_guides<<<mt_blocks(ngm + ngc, 256), 256, 0, ctx->stream>>>(e->marg, e->cond, w, h, e->gmarg, e->gcond);
    // tile-region tables for the light-grid guided sampler (16 x 32 tiles)
    e->ttab = nullptr;
    if (w >= LG_RX && h >= LG_RY) {
        e->tsxm = w / LG_RX + 2;
        e->tsym = h / LG_RY + 2;
        e->tstride = env_tile_stride(e->tsxm, e->tsym);
        const int64_t nt = (int64_t)LG_RX * LG_RY * e->tstride;
        if (nt <= (int64_t)1 << 26 && hipMalloc(&e->ttab, sizeof(float) * (size_t)nt) == hipSuccess) {
            EnvView E = env_view(e);
            E.ttab = nullptr; // entries are evaluated from the SAT
            k_env_tile_tables<<<mt_blocks(nt, 256), 256, 0, ctx->stream>>>(E, e->ttab);
        } else {
            e->ttab = nullptr;
        }
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        (void)hipFree(e->tab);
        if (e->ttab) (void)hipFree(e->ttab);
        delete e;
        mt_set_error("mt_env_create: %s", hipGetErrorString(err));
        return MT_ERR_HIP;
    }
    *out = e;
    return MT_OK;
}

extern "C" int32_t mt_env_destroy(mt_envmap* e)
{
    if (!e) return MT_OK;
    MtFirstError fe;
    fe(hipStreamSynchronize(e->ctx->stream));
    fe(hipFree(e->tab));
    if (e->ttab) fe(hipFree(e->ttab));
    delete e;
    return fe.status("mt_env_destroy");
}

extern "C" int32_t mt_sat_sample_region(mt_context* ctx, const float* func, const float* sat, int32_t nu,
                                        int32_t nv, int32_t minx, int32_t miny, int32_t sx, int32_t sy,
                                        int32_t n, const float* u, float* uv, int32_t* offset, float* pdf)
{
    MT_ARG_CHECK(ctx && func && sat && u && uv && offset && pdf, "mt_sat_sample_region: NULL argument");
    MT_ARG_CHECK(minx >= 0 && miny >= 0 && sx > 0 && sy > 0 && minx + sx <= nu && miny + sy <= nv,
                 "mt_sat_sample_region: region outside the map");
    if (n <= 0) return MT_OK;
    SatRegion R{nu, nv, minx, miny, sx, sy};
    k_region_sample<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(func, sat, R, n,
                                                                reinterpret_cast<const float2*>(u),
                                                                reinterpret_cast<float2*>(uv),
                                                                reinterpret_cast<int2*>(offset), pdf);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_env_illuminate(mt_context* ctx, const mt_envmap* env, int32_t n, const float* u,
                                     float* dir, float* contrib, float* pdfw, int32_t* texel)
{
    MT_ARG_CHECK(ctx && env && u && dir && contrib && pdfw && texel, "mt_env_illuminate: NULL argument");
    if (n <= 0) return MT_OK;
    EnvView E = env_view(env);
    k_env_illuminate<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(E, n, reinterpret_cast<const float2*>(u),
                                                                 reinterpret_cast<float4*>(dir),
                                                                 reinterpret_cast<float4*>(contrib), pdfw,
                                                                 texel);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_env_radiance(mt_context* ctx, const mt_envmap* env, int32_t n, const float* dir,
                                   float* contrib, float* pdfw, int32_t* texel)
{
    MT_ARG_CHECK(ctx && env && dir && contrib && pdfw && texel, "mt_env_radiance: NULL argument");
    if (n <= 0) return MT_OK;
    EnvView E = env_view(env);
    k_env_radiance<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(E, n, reinterpret_cast<const float4*>(dir),
                                                               reinterpret_cast<float4*>(contrib), pdfw,
                                                               texel);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_lightgrid_set_camera(mt_lightgrid* grid, const float* pos)
{
    MT_ARG_CHECK(grid && pos, "mt_lightgrid_set_camera: NULL argument");
    grid->cam[0] = pos[0];
    grid->cam[1] = pos[1];
    grid->cam[2] = pos[2];
    return MT_OK;
}

// Tiles of the grid map onto texel rectangles of the environment map; every
// rectangle must hold a texel (setRegion, SAT.cl:12-24, leaves an empty
// region undefined), so the map needs at least one texel per tile column/row.
static int32_t check_adaptive_env(const mt_envmap* env, const char* fn)
{
    MT_ARG_CHECK(env->w >= LG_RX && env->h >= LG_RY, "%s: environment map %d x %d smaller than the %d x %d tiles",
                 fn, env->w, env->h, LG_RX, LG_RY);
    return MT_OK;
}

extern "C" int32_t mt_env_illuminate_adaptive(mt_context* ctx, const mt_envmap* env, const mt_lightgrid* grid,
                                              int32_t n, const float* hit_p, const float* u, float* dir,
                                              float* contrib, float* pdfw, float* lum, int32_t* grid_index)
{
    MT_ARG_CHECK(ctx && env && grid && hit_p && u && dir && contrib && pdfw,
                 "mt_env_illuminate_adaptive: NULL argument");
    MT_ARG_CHECK(n >= 0, "mt_env_illuminate_adaptive: negative n");
    int32_t rc = check_adaptive_env(env, "mt_env_illuminate_adaptive");
    if (rc) return rc;
    if (n == 0) return MT_OK;
    k_env_illuminate_adaptive<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(
        env_view(env), grid->func, grid->sat, grid->tab_valid ? grid->tab : nullptr, grid->cam[0], grid->cam[1],
        grid->cam[2], n,
        reinterpret_cast<const float4*>(hit_p), reinterpret_cast<const float4*>(u), reinterpret_cast<float4*>(dir),
        reinterpret_cast<float4*>(contrib), pdfw, lum, grid_index);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_env_radiance_adaptive(mt_context* ctx, const mt_envmap* env, mt_lightgrid* grid, int32_t n,
                                            const float* dir, const float* hit_p, int32_t accumulate,
                                            float* contrib, float* pdfa, float* lum, int32_t* grid_index)
{
    MT_ARG_CHECK(ctx && env && grid && dir && hit_p, "mt_env_radiance_adaptive: NULL argument");
    MT_ARG_CHECK(n >= 0, "mt_env_radiance_adaptive: negative n");
    int32_t rc = check_adaptive_env(env, "mt_env_radiance_adaptive");
    if (rc) return rc;
    if (n == 0) return MT_OK;
    k_env_radiance_adaptive<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(
        env_view(env), accumulate ? grid->accum : nullptr, grid->cam[0], grid->cam[1], grid->cam[2], n,
        reinterpret_cast<const float4*>(dir), reinterpret_cast<const float4*>(hit_p),
        reinterpret_cast<float4*>(contrib), pdfa, lum, grid_index);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_env_attach_lightgrid(mt_envmap* env, mt_lightgrid* grid)
{
    MT_ARG_CHECK(env != nullptr, "mt_env_attach_lightgrid: NULL environment");
    if (grid) {
        int32_t rc = check_adaptive_env(env, "mt_env_attach_lightgrid");
        if (rc) return rc;
    }
    env->grid = grid;
    return MT_OK;
}

// Folds the render loop's fixed-point accumulation into the grid's accum
// (stream-ordered; no allocation, so a captured render may include it).
int32_t mt_lightgrid_fold(hipStream_t st, mt_lightgrid* g)
{
    k_lg_fold<<<mt_blocks(LG_SIZE, 256), 256, 0, st>>>(g->pend, g->accum);
    MT_LAUNCH_CHECK();
    return MT_OK;
}
