// Stage C device functions: light-grid index math and subgrid sampling
// (src/cl/kernel/LightGrid.cl:39-343), region SAT sampling (SAT.cl:1-342),
// spherical mappings (Util.cl:397-439) and environment light
// illuminate/radiance (Light.cl:129-137,195-227,270-286).
//
// All comparisons and divisions follow the reference order so cell, tile and
// texel picks round identically to the CPU restatement.  Deviations: reads
// the reference makes outside its arrays (getSubgridValueSAT evaluates
// sat[index] before select() discards it, LightGrid.cl:145; getFunc with
// index -1, SAT.cl:95-99; texel y == height, Util.cl:406) return 0 or are
// clamped to the map here.
#pragma once

#include "mt_devmath.h"
#include "mt_internal.h"

// ---------------------------------------------------------------- CDF search
// upperBound (LightGrid.cl:212-247, SAT.cl:206-242) by bisection: the number
// of leading entries of a[b, e) with !(u < a[i]) when a is non-decreasing,
// and for any a the index the reference's bisection returns.  Each round
// loads the range's midpoint and the midpoints of both halves together and
// takes two bisection steps (the second compares against the half the first
// step keeps): half the dependent loads of one step per round, for three
// loads per two steps.  Every comparison is the bisection's own and every
// load lies inside [b, e).  (Four steps per round, 15 loads, was measured
// slower: the shading kernels that call this are bound by VALU issue, and
// the extra selects cost more than the saved latency.)
__device__ __forceinline__ int cdf_search2(const float* __restrict__ a, int b, int e, float u)
{
    while (b != e) {
        const int mid = b + (e - b) / 2;
        const int lm = b + (mid - b) / 2;               // midpoint of [b, mid)
        const int rm = (mid + 1) + (e - (mid + 1)) / 2; // midpoint of [mid + 1, e)
        const float vm = a[mid];
        const float vl = b < mid ? a[lm] : 0.0f;
        const float vr = mid + 1 < e ? a[rm] : 0.0f;
        const bool right = !(u < vm);
        b = right ? mid + 1 : b;
        e = right ? e : mid;
        if (b != e) {
            const int m2 = b + (e - b) / 2; // rm or lm
            const bool r2 = !(u < (right ? vr : vl));
            b = r2 ? m2 + 1 : b;
            e = r2 ? e : m2;
        }
    }
    return b;
}

// ---------------------------------------------------------------- light grid
__device__ __forceinline__ int lg_global_index(int s, int lx, int ly)
{ // globalIndexInSubgrid, LightGrid.cl:104-139
    int cx = s % LG_CX;
    int cy = s / LG_CX;
    int gi = cx * LG_RX + cy * (LG_RX * LG_RY) * LG_CX;
    int gx = gi % LG_NU;
    int gy = gi / LG_NV; // subgridGlobalIndexXY divides by nv (:122); nu == nv
    return (gx + lx) + (gy + ly) * LG_NU;
}

__device__ __forceinline__ float lg_sat(const float* __restrict__ sat, int s, int sx, int sy)
{ // getSubgridValueSAT, :142-148
    bool in = !(sx >= LG_RX || sx < 0) && !(sy >= LG_RY || sy < 0);
    return in ? sat[lg_global_index(s, sx, sy)] : 0.0f;
}

__device__ __forceinline__ float lg_range(const float* __restrict__ sat, int s, int minx, int miny,
                                          int maxx, int maxy)
{ // getSubgridSATRange, :150-157
    float a1 = lg_sat(sat, s, maxx, maxy);
    float b1 = lg_sat(sat, s, minx - 1, maxy);
    float c1 = lg_sat(sat, s, maxx, miny - 1);
    float d1 = lg_sat(sat, s, minx - 1, miny - 1);
    return ((a1 + d1) - b1) - c1;
}

__device__ __forceinline__ float lg_func_int(const float* sat, int s, int y)
{
    return lg_range(sat, s, 0, y, LG_RX - 1, y);
}

__device__ __forceinline__ float lg_cond(const float* sat, int s, int x, int y)
{ // getSubgridConditional, :183-187
    float fi = lg_func_int(sat, s, y);
    return lg_range(sat, s, 0, y, x - 1, y) / fi;
}

__device__ __forceinline__ float lg_marg(const float* sat, int s, int y)
{ // getSubgridMarginal, :189-193
    float last = lg_sat(sat, s, LG_RX - 1, LG_RY - 1);
    return lg_sat(sat, s, LG_RX - 1, y - 1) / last;
}

// sampleSubgridContinuous, LightGrid.cl:310-320 (marginal :279-296, conditional :249-266,
// branchless upper bounds :212-247)
__device__ __forceinline__ void lg_sample(const float* __restrict__ func, const float* __restrict__ sat,
                                          int s, float u0, float u1, int& ox, int& oy, float& uu,
                                          float& vv, float& pdf)
{
    int b = 0, e = LG_RY;
    while (b != e) {
        int mid = b + (e - b) / 2;
        bool right = !(u1 < lg_marg(sat, s, mid));
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    oy = b - 1 > 0 ? b - 1 : 0;
    float m0 = lg_marg(sat, s, oy);
    float du = (u1 - m0) / (lg_marg(sat, s, oy + 1) - m0);
    float last = lg_sat(sat, s, LG_RX - 1, LG_RY - 1);
    float pdf_v = (lg_func_int(sat, s, oy) * (float)(LG_RX * LG_RY)) / last;
    vv = ((float)oy + du) / (float)LG_RY;
    b = 0;
    e = LG_RX;
    while (b != e) {
        int mid = b + (e - b) / 2;
        bool right = !(u0 < lg_cond(sat, s, mid, oy));
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    ox = b - 1 > 0 ? b - 1 : 0;
    float c0 = lg_cond(sat, s, ox, oy);
    float du2 = (u0 - c0) / (lg_cond(sat, s, ox + 1, oy) - c0);
    float pdf_u = func[lg_global_index(s, ox, oy)] / lg_func_int(sat, s, oy);
    uu = ((float)ox + du2) / (float)LG_RX;
    pdf = pdf_u * pdf_v;
}

// Per-cell sampling tables: every quantity lg_sample evaluates from SAT
// corners, each entry the same float expression computed once per SAT build,
// laid out cell-major so one sample reads a few contiguous cache lines
// instead of ~60 scattered corners.  Cell s owns LG_TAB floats:
//   [0, 33)       marg[y]  = getSubgridMarginal(y),        y in [0, 32]
//   [33, 65)      fint[y]  = getSubgridFuncIntConditional(y)
//   [65]          last     = getSubgridValueSAT(15, 31)
//   [80, 80+544)  cond[y][x] = getSubgridConditional(x, y), x in [0, 16], row stride 17
constexpr int LG_TAB = 640;
constexpr int LG_TAB_FINT = 33, LG_TAB_LAST = 65, LG_TAB_COND = 80, LG_TAB_CSTRIDE = LG_RX + 1;

__device__ __forceinline__ float lg_table_entry(const float* __restrict__ sat, int s, int i)
{
    if (i < LG_TAB_FINT) return lg_marg(sat, s, i);
    if (i < LG_TAB_LAST) return lg_func_int(sat, s, i - LG_TAB_FINT);
    if (i == LG_TAB_LAST) return lg_sat(sat, s, LG_RX - 1, LG_RY - 1);
    if (i < LG_TAB_COND || i >= LG_TAB_COND + LG_RY * LG_TAB_CSTRIDE) return 0.0f;
    const int k = i - LG_TAB_COND;
    return lg_cond(sat, s, k % LG_TAB_CSTRIDE, k / LG_TAB_CSTRIDE);
}

// lg_sample reading the tables: the same bisections over the same values
// and the same arithmetic on them, so every output has the same bits.
__device__ __forceinline__ void lg_sample_tab(const float* __restrict__ func, const float* __restrict__ tab,
                                              int s, float u0, float u1, int& ox, int& oy, float& uu, float& vv,
                                              float& pdf)
{
    const float* T = tab + (size_t)s * LG_TAB;
    int b = cdf_search2(T, 0, LG_RY, u1);
    oy = b - 1 > 0 ? b - 1 : 0;
    const float m0 = T[oy];
    const float du = (u1 - m0) / (T[oy + 1] - m0);
    const float fi = T[LG_TAB_FINT + oy];
    const float pdf_v = (fi * (float)(LG_RX * LG_RY)) / T[LG_TAB_LAST];
    vv = ((float)oy + du) / (float)LG_RY;
    const float* C = T + LG_TAB_COND + oy * LG_TAB_CSTRIDE;
    b = cdf_search2(C, 0, LG_RX, u0);
    ox = b - 1 > 0 ? b - 1 : 0;
    const float c0 = C[ox];
    const float du2 = (u0 - c0) / (C[ox + 1] - c0);
    const float pdf_u = func[lg_global_index(s, ox, oy)] / fi;
    uu = ((float)ox + du2) / (float)LG_RX;
    pdf = pdf_u * pdf_v;
}

// ---------------------------------------------------------------- spherical maps
// The two angles getSphericalGridXY bins (Util.cl:399-400); callers that bin
// one direction on several grids compute them once (same values, same bits).
struct SphAngles {
    float phi, theta;
};
__device__ __forceinline__ SphAngles sph_angles(v3 d) { return SphAngles{mt_xm_acosf(d.y), mt_xm_atan2f(d.z, d.x)}; }

__device__ __forceinline__ void sph_xy_of(const SphAngles& a, int w, int h, int& x, int& y)
{ // getSphericalGridXY, Util.cl:397-410
    x = (int)((0.5f - (0.5f * a.theta) / MT_PI) * (float)w);
    y = (int)((a.phi / MT_PI) * (float)h);
}

__device__ __forceinline__ int sph_index_of(const SphAngles& a, int w, int h)
{ // getSphericalGridIndex, Util.cl:413-417 (clamped to the map)
    int x, y;
    sph_xy_of(a, w, h, x, y);
    int idx = (int)((float)x + (float)y * (float)w);
    idx = idx < 0 ? 0 : idx;
    return idx > w * h - 1 ? w * h - 1 : idx;
}

__device__ __forceinline__ void sph_xy(int w, int h, v3 d, int& x, int& y) { sph_xy_of(sph_angles(d), w, h, x, y); }

__device__ __forceinline__ int sph_index(int w, int h, v3 d) { return sph_index_of(sph_angles(d), w, h); }

__device__ __forceinline__ v3 sph_direction(float u, float v)
{ // getSphericalDirection, Util.cl:423-439
    float phi = MT_PI * v;
    float theta = (u * 2.0f) * MT_PI;
    return mk3(-mt_xm_sinf(phi) * mt_xm_cosf(theta), mt_xm_cosf(phi), mt_xm_sinf(phi) * mt_xm_sinf(theta));
}

// ---------------------------------------------------------------- region SAT
struct SatRegion {
    int nu, nv, minx, miny, sx, sy;
};

__device__ __forceinline__ float r_value(const SatRegion& R, const float* __restrict__ sat, int x, int y)
{ // getValueSAT, SAT.cl:135-147
    if (x < 0 || y < 0) return 0.0f;
    if (x > R.minx + R.sx - 1) return 0.0f;
    if (y > R.miny + R.sy - 1) return 0.0f;
    if (x > R.nu - 1 || y > R.nv - 1) return 0.0f;
    return sat[(size_t)y * R.nu + x];
}

__device__ __forceinline__ float r_range(const SatRegion& R, const float* sat, int minx, int miny,
                                         int maxx, int maxy)
{ // getSATRange, SAT.cl:149-156
    float d1 = r_value(R, sat, maxx, maxy);
    float c1 = r_value(R, sat, minx - 1, maxy);
    float b1 = r_value(R, sat, maxx, miny - 1);
    float a1 = r_value(R, sat, minx - 1, miny - 1);
    return ((a1 + d1) - b1) - c1;
}

__device__ __forceinline__ float r_func_int(const SatRegion& R, const float* sat, int y)
{
    return r_range(R, sat, R.minx, y, R.minx + R.sx - 1, y);
}
__device__ __forceinline__ float r_cond(const SatRegion& R, const float* sat, int x, int y)
{
    float fi = r_func_int(R, sat, y);
    return r_range(R, sat, R.minx, y, x - 1, y) / fi;
}
__device__ __forceinline__ float r_marg(const SatRegion& R, const float* sat, int y)
{
    float last = r_range(R, sat, R.minx, R.miny, R.minx + R.sx - 1, R.miny + R.sy - 1);
    return r_range(R, sat, R.minx, R.miny, R.minx + R.sx - 1, y - 1) / last;
}
__device__ __forceinline__ float r_func(const SatRegion& R, const float* __restrict__ func, int x, int y)
{
    if (x < 0 || y < 0 || x > R.nu - 1 || y > R.nv - 1) return 0.0f;
    return func[(size_t)y * R.nu + x];
}
__device__ __forceinline__ float r_pdf_cond(const SatRegion& R, const float* sat, const float* func,
                                            int x, int y)
{
    return r_func(R, func, x, y) / r_func_int(R, sat, y);
}
__device__ __forceinline__ float r_pdf_marg(const SatRegion& R, const float* sat, int y)
{
    float fi = r_func_int(R, sat, y);
    float last = r_range(R, sat, R.minx, R.miny, R.minx + R.sx - 1, R.miny + R.sy - 1);
    return (fi * (float)(R.sx * R.sy)) / last;
}

// sampleContinuous, SAT.cl:244-323
__device__ __forceinline__ void r_sample(const SatRegion& R, const float* __restrict__ sat,
                                         const float* __restrict__ func, float u0, float u1,
                                         float& uu, float& vv, int& ox, int& oy, float& pdf)
{
    int b = R.miny, e = R.miny + R.sy;
    while (b != e) {
        int mid = b + (e - b) / 2;
        bool right = !(u1 < r_marg(R, sat, mid));
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    oy = b - 1 > 0 ? b - 1 : 0;
    float m0 = r_marg(R, sat, oy);
    float du = (u1 - m0) / (r_marg(R, sat, oy + 1) - m0);
    float pdf_v = r_pdf_marg(R, sat, oy);
    vv = ((float)oy + du) / (float)R.nv;
    b = R.minx;
    e = R.minx + R.sx;
    while (b != e) {
        int mid = b + (e - b) / 2;
        bool right = !(u0 < r_cond(R, sat, mid, oy));
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    ox = b - 1 > 0 ? b - 1 : 0;
    float c0 = r_cond(R, sat, ox, oy);
    float du2 = (u0 - c0) / (r_cond(R, sat, ox + 1, oy) - c0);
    float pdf_u = r_pdf_cond(R, sat, func, ox, oy);
    uu = ((float)ox + du2) / (float)R.nu;
    pdf = pdf_u * pdf_v;
}

// ---------------------------------------------------------------- environment light
// Optional exact tables of the full-map quantities the sampler evaluates
// from SAT corners (each entry is the same float expression, computed once):
//   marg[y]   = getMarginal(y),                 y in [0, h]
//   cond[y][x]= getConditional(x, y),           x in [0, w], row stride w+1
//   fint[y]   = getFuncIntConditional(y),       y in [0, h]
//   pmarg[y]  = getPdfContinuousMarginal(y),    y in [0, h]
struct EnvView {
    const float4* rgba;
    const float* lum;
    const float* sat;
    int w, h;
    const float* marg;  // null: evaluate from the SAT
    const float* cond;
    const float* fint;
    const float* pmarg;
    const int* gmarg;   // search guides (null: plain binary search)
    const int* gcond;
    const float* ttab;  // light-grid tile-region tables (null: evaluate from the SAT)
    int tsxm, tsym, tstride;
};

inline EnvView env_view(const mt_envmap* env)
{
    if (!env)
        return EnvView{nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, 0, 0, 0};
    return EnvView{reinterpret_cast<const float4*>(env->rgba), env->lum, env->sat, env->w, env->h,
                   env->marg, env->cond, env->fint, env->pmarg, env->gmarg, env->gcond,
                   env->ttab, env->tsxm, env->tsym, env->tstride};
}

// upperBoundMarginal / upperBoundConditional (SAT.cl:206-242): the number of
// leading entries of a[b, e) with !(u < a[i]), by bisection.
__device__ __forceinline__ int cdf_search(const float* __restrict__ a, int b, int e, float u)
{
    while (b != e) {
        int mid = b + (e - b) / 2;
        bool right = !(u < a[mid]);
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    return b;
}

// The same search over a[0, n) narrowed by a guide: g[k] = cdf_search(a, 0,
// n, k / m) for k in [0, m+1] (m a power of two, so u * m and k / m are
// exact).  The predicate !(u < a[i]) is monotone in u, and in i for a
// non-decreasing CDF, so for k = floor(u * m) the answer lies in
// [g[k], g[k+1]] and bisecting that range returns the full search's index.
__device__ __forceinline__ int cdf_search_guided(const float* __restrict__ a, int n, const int* __restrict__ g,
                                                 int m, float u)
{
    const float s = u * (float)m;
    if (!(s >= 0.0f && s <= (float)m)) return cdf_search(a, 0, n, u);
    const int k = (int)s;
    const int lo = g[k], hi = g[k + 1];
    // a table that is not non-decreasing can give an inverted bracket:
    // search the whole table instead
    if (!(0 <= lo && lo <= hi && hi <= n)) return cdf_search(a, 0, n, u);
    return cdf_search(a, lo, hi, u);
}

__device__ __forceinline__ void env_sample_tables(const EnvView& E, float u0, float u1, float& uu,
                                                  float& vv, int& ox, int& oy, float& pdf)
{ // sampleContinuous (SAT.cl:244-323) reading the tables above
    int b = E.gmarg ? cdf_search_guided(E.marg, E.h, E.gmarg, ENV_GUIDE_M, u1) : cdf_search(E.marg, 0, E.h, u1);
    oy = b - 1 > 0 ? b - 1 : 0;
    float m0 = E.marg[oy];
    float du = (u1 - m0) / (E.marg[oy + 1] - m0);
    float pdf_v = E.pmarg[oy];
    vv = ((float)oy + du) / (float)E.h;
    const float* C = E.cond + (size_t)oy * (E.w + 1);
    b = E.gcond ? cdf_search_guided(C, E.w, E.gcond + (size_t)oy * (ENV_GUIDE_C + 2), ENV_GUIDE_C, u0)
                : cdf_search(C, 0, E.w, u0);
    ox = b - 1 > 0 ? b - 1 : 0;
    float c0 = C[ox];
    float du2 = (u0 - c0) / (C[ox + 1] - c0);
    float pdf_u = E.lum[(size_t)oy * E.w + ox] / E.fint[oy];
    uu = ((float)ox + du2) / (float)E.w;
    pdf = pdf_u * pdf_v;
}

// illuminateEnvironmentLight, Light.cl:195-227: direction, radiance, solid-angle pdf
__device__ __forceinline__ float4 env_illuminate(const EnvView& E, float u0, float u1, v3& dir,
                                                 float& pdfw, int& texel)
{
    SatRegion R{E.w, E.h, 0, 0, E.w, E.h};
    float uu, vv, pdf;
    int ox, oy;
    if (E.marg) env_sample_tables(E, u0, u1, uu, vv, ox, oy, pdf);
    else r_sample(R, E.sat, E.lum, u0, u1, uu, vv, ox, oy, pdf);
    dir = sph_direction(uu, vv);
    texel = sph_index(E.w, E.h, dir);
    float s = mt_xm_sinf(vv * MT_PI);
    pdfw = pdf / (((2.0f * MT_PI) * MT_PI) * mt_xm_sinf(vv * MT_PI));
    if (s == 0.0f) pdfw = 0.0f;
    return E.rgba[texel];
}

// getRadianceEnvironmentLight, Light.cl:270-286
__device__ __forceinline__ float4 env_radiance(const EnvView& E, v3 d, float& pdfw, int& texel)
{
    SatRegion R{E.w, E.h, 0, 0, E.w, E.h};
    const SphAngles a = sph_angles(d);
    int x, y;
    sph_xy_of(a, E.w, E.h, x, y);
    texel = sph_index_of(a, E.w, E.h);
    float p;
    if (E.marg && x >= 0 && y >= 0 && x <= E.w && y <= E.h) // getFunc outside the map reads 0
        p = (r_func(R, E.lum, x, y) / E.fint[y]) * E.pmarg[y];
    else
        p = r_pdf_cond(R, E.sat, E.lum, x, y) * r_pdf_marg(R, E.sat, y);
    float v = (float)y / (float)E.h;
    pdfw = p / (((2.0f * MT_PI) * MT_PI) * mt_xm_sinf(v * MT_PI));
    return E.rgba[texel];
}

// ---------------------------------------------------------------- adaptive environment light
// subgridIndexFromCamera, LightGrid.cl:68-78: the 100 x 50 cell of the
// direction from the grid's camera position to the hit point.
__device__ __forceinline__ int lg_cell_from_camera(v3 cam, v3 hit)
{
    return sph_index(LG_CX, LG_CY, norm3(sub3(hit, cam)));
}

// getSubgridUnitBound (LightGrid.cl:158-167) followed by
// getSubRegionFromUnitBound (SAT.cl:28-44): the env-map texel rectangle of
// tile (tx, ty) of a cell, corners truncated from float products.
__device__ __forceinline__ SatRegion lg_tile_region(int w, int h, int tx, int ty)
{
    const float ux0 = (float)tx / (float)LG_RX, uy0 = (float)ty / (float)LG_RY;
    const float ux1 = (float)(tx + 1) / (float)LG_RX, uy1 = (float)(ty + 1) / (float)LG_RY;
    const int minx = (int)((float)w * ux0), miny = (int)((float)h * uy0);
    const int maxx = (int)((float)w * ux1), maxy = (int)((float)h * uy1);
    return SatRegion{w, h, minx, miny, maxx - minx, maxy - miny};
}

// Tile-region tables.  For tile (tx, ty) with region R = lg_tile_region(...),
// every quantity r_sample / getPdfSAT evaluates from SAT corners inside
// the rows and columns the sampler can reach (the bisections stay in the
// region; the offsets max(0, ptr - 1) reach one row / column before it):
//   marg[j] = r_marg(R, miny - 1 + j),                j in [0, sym + 2)
//   fint[j] = r_func_int(R, miny - 1 + j),            j in [0, sym + 1)
//   last    = the region's total
//   cond[j][i] = r_cond(R, minx - 1 + i, miny - 1 + j), i in [0, sxm + 2)
// (entries for negative rows / columns, and past the region, are 0 and
// never read).  sxm, sym bound every tile's width and height.
__host__ __device__ inline int env_tile_stride(int sxm, int sym)
{
    const int t = (sym + 2) + (sym + 1) + 1 + (sym + 1) * (sxm + 2);
    return (t + 31) & ~31;
}

__device__ __forceinline__ float env_tile_entry(const EnvView& E, int k, int i)
{
    const SatRegion R = lg_tile_region(E.w, E.h, k % LG_RX, k / LG_RX);
    const int nm = E.tsym + 2, nf = E.tsym + 1, cs = E.tsxm + 2;
    if (i < nm) {
        const int y = R.miny - 1 + i;
        return (y >= 0 && y <= R.miny + R.sy) ? r_marg(R, E.sat, y) : 0.0f;
    }
    i -= nm;
    if (i < nf) {
        const int y = R.miny - 1 + i;
        return (y >= 0 && y < R.miny + R.sy) ? r_func_int(R, E.sat, y) : 0.0f;
    }
    i -= nf;
    if (i == 0) return r_range(R, E.sat, R.minx, R.miny, R.minx + R.sx - 1, R.miny + R.sy - 1);
    i -= 1;
    if (i < nf * cs) {
        const int y = R.miny - 1 + i / cs, x = R.minx - 1 + i % cs;
        return (y >= 0 && x >= 0 && y < R.miny + R.sy && x <= R.minx + R.sx) ? r_cond(R, E.sat, x, y) : 0.0f;
    }
    return 0.0f;
}

// r_sample on tile region R reading its table T: the same bisections over
// the same values, the same arithmetic on them.
__device__ __forceinline__ void r_sample_tile(const EnvView& E, const SatRegion& R, const float* __restrict__ T,
                                              float u0, float u1, float& uu, float& vv, int& ox, int& oy, float& pdf)
{
    const int nm = E.tsym + 2, nf = E.tsym + 1, cs = E.tsxm + 2;
    const float* M = T - (R.miny - 1);        // M[y] = marg of row y
    const float* FI = T + nm - (R.miny - 1);  // FI[y]
    const float last = T[nm + nf];
    int b = cdf_search2(M, R.miny, R.miny + R.sy, u1);
    oy = b - 1 > 0 ? b - 1 : 0;
    const float m0 = M[oy];
    const float du = (u1 - m0) / (M[oy + 1] - m0);
    const float fi = FI[oy];
    const float pdf_v = (fi * (float)(R.sx * R.sy)) / last;
    vv = ((float)oy + du) / (float)R.nv;
    const float* C = T + nm + nf + 1 + (oy - (R.miny - 1)) * cs - (R.minx - 1); // C[x] = cond(x, oy)
    b = cdf_search2(C, R.minx, R.minx + R.sx, u0);
    ox = b - 1 > 0 ? b - 1 : 0;
    const float c0 = C[ox];
    const float du2 = (u0 - c0) / (C[ox + 1] - c0);
    const float pdf_u = r_func(R, E.lum, ox, oy) / fi;
    uu = ((float)ox + du2) / (float)R.nu;
    pdf = pdf_u * pdf_v;
}

// illuminateEnvironmentLight1, Light.cl:140-193: tile from the cell's
// distribution, texel from the tile's SAT region.  The returned pdf is the
// in-tile pdf only (the tile pick's pdf is not applied, :163,182); tile_pdf
// is that pick's continuous pdf over the cell's 16 x 32 tiles (pdfLL, which
// the reference computes and leaves unused).
__device__ __forceinline__ float4 env_illuminate_adaptive(const EnvView& E, const float* __restrict__ lg_func,
                                                          const float* __restrict__ lg_sat_arr,
                                                          const float* __restrict__ lg_tab, v3 cam, v3 hit,
                                                          float4 u, v3& dir, float& pdfw, float& lum, int& gidx,
                                                          float& tile_pdf)
{
    const int s = lg_cell_from_camera(cam, hit);
    int lx, ly;
    float luu, lvv, lpdf;
    if (lg_tab) lg_sample_tab(lg_func, lg_tab, s, u.x, u.y, lx, ly, luu, lvv, lpdf);
    else lg_sample(lg_func, lg_sat_arr, s, u.x, u.y, lx, ly, luu, lvv, lpdf);
    const SatRegion R = lg_tile_region(E.w, E.h, lx, ly);
    float uu, vv, pdf;
    int ox, oy;
    if (E.ttab) r_sample_tile(E, R, E.ttab + (size_t)(ly * LG_RX + lx) * E.tstride, u.z, u.w, uu, vv, ox, oy, pdf);
    else r_sample(R, E.sat, E.lum, u.z, u.w, uu, vv, ox, oy, pdf);
    dir = sph_direction(uu, vv);
    const float4 c = E.rgba[sph_index(E.w, E.h, dir)];
    const float sn = mt_xm_sinf(vv * MT_PI);
    pdfw = pdf / (((2.0f * MT_PI) * MT_PI) * mt_xm_sinf(vv * MT_PI));
    if (sn == 0.0f) pdfw = 0.0f;
    lum = luminance3(c.x, c.y, c.z);
    gidx = lg_global_index(s, lx, ly);
    tile_pdf = lpdf;
    return c;
}

// getRadianceEnvironmentLight1, Light.cl:229-268, without the accumulation
// (the caller adds lum into accum[gidx], accumLightGrid :35-38); cell and
// tile are the grid cell of the hit point and the tile (ty * 16 + tx) of
// the direction.
__device__ __forceinline__ float4 env_radiance_adaptive(const EnvView& E, v3 cam, v3 d, v3 hit, float& pdfa,
                                                        float& lum, int& gidx, int& cell, int& tile)
{
    const int s = lg_cell_from_camera(cam, hit);
    const SphAngles a = sph_angles(d);
    const int t = sph_index_of(a, LG_RX, LG_RY); // tileIndexXYFromCamera, LightGrid.cl:80-90
    const int tx = t % LG_RX, ty = t / LG_RX;
    const SatRegion R = lg_tile_region(E.w, E.h, tx, ty);
    int x, y;
    sph_xy_of(a, E.w, E.h, x, y);
    float p;
    if (E.ttab && y >= R.miny - 1 && y >= 0 && y < R.miny + R.sy) { // the table's rows: same values
        const float* T = E.ttab + (size_t)t * E.tstride;
        const float fi = T[E.tsym + 2 + (y - (R.miny - 1))];
        const float last = T[(E.tsym + 2) + (E.tsym + 1)];
        p = (r_func(R, E.lum, x, y) / fi) * ((fi * (float)(R.sx * R.sy)) / last); // getPdfSAT, SAT.cl:336-342
    } else {
        p = r_pdf_cond(R, E.sat, E.lum, x, y) * r_pdf_marg(R, E.sat, y);
    }
    const float sv = (float)y / (float)E.h;
    pdfa = p / (((2.0f * MT_PI) * MT_PI) * mt_xm_sinf(sv * MT_PI));
    const float4 c = E.rgba[sph_index_of(a, E.w, E.h)];
    lum = luminance3(c.x, c.y, c.z);
    gidx = lg_global_index(s, tx, ty);
    cell = s;
    tile = t;
    return c;
}

// ---------------------------------------------------------------- light grid in the bounce loop
// What the bounce loop's guided environment mode reads: the grid's func,
// SAT and sampling tables (tab null: evaluate from the SAT), the grid's
// camera position, and the fixed-point accumulation it adds into.
// mode 0: full-map environment sampling (illuminateEnvironmentLight /
// getRadianceEnvironmentLight); 1: the guided pair with the reference's pdf
// (in-tile only); 2: the guided pair with the tile pick's pdf applied.
struct LgView {
    const float* func;
    const float* sat;
    const float* tab;
    float cx, cy, cz;
    unsigned long long* pend;
    int mode;
};

// Continuous pdf of tile (tx, ty) in cell s: the value sampleSubgridContinuous
// returns when it picks that tile, (func / rowInt) * (rowInt * 512 / total);
// 0 for a tile the cell cannot pick (row or total 0, or not finite).
__device__ __forceinline__ float lg_tile_pdf(const LgView& G, int s, int tx, int ty)
{
    float fi, last;
    if (G.tab) {
        const float* T = G.tab + (size_t)s * LG_TAB;
        fi = T[LG_TAB_FINT + ty];
        last = T[LG_TAB_LAST];
    } else {
        fi = lg_func_int(G.sat, s, ty);
        last = lg_sat(G.sat, s, LG_RX - 1, LG_RY - 1);
    }
    const float f = G.func[lg_global_index(s, tx, ty)];
    const float p = (f / fi) * ((fi * (float)(LG_RX * LG_RY)) / last);
    return (fi > 0.0f && last > 0.0f && p > 0.0f && p < INFINITY) ? p : 0.0f;
}

// Fixed-point addend of a luminance (units of 2^-20, round half up); 0 for
// anything not finite and non-negative.
__host__ __device__ inline unsigned long long lg_fixed(float lum)
{
    const double v = (double)lum * 1048576.0;
    return (v >= 0.0 && v < 9.0e15) ? (unsigned long long)floor(v + 0.5) : 0ull;
}

// accumLightGrid (Light.cl:35-38) for the lanes with `add`: entry gidx of
// the fixed-point accumulation gains q (in the work-group's copy of it).
// Call with the whole wave active.
// Lanes that share an entry with the wave's first pending lane are summed
// and added with one atomic (camera-coherent misses share a few entries);
// after four such rounds, or as soon as a round finds at most two lanes on
// its entry (scattered entries, as for paths escaping after a bounce: the
// shuffle sum would save one atomic), the rest add their own.  Integer
// sums: the result does not depend on how the additions are grouped.
__device__ __forceinline__ void lg_wave_accum(bool add, int gidx, unsigned long long q,
                                              unsigned long long* __restrict__ pend)
{
    pend += (size_t)(blockIdx.x % LG_PEND_COPIES) * LG_SIZE; // this work-group's copy
    add = add && q != 0ull;
    uint64_t m = __ballot(add);
    for (int round = 0; round < 4 && m; round++) {
        const int leader = __builtin_ctzll(m);
        const int g = __shfl(gidx, leader);
        const bool mine = add && gidx == g;
        if (__popcll(__ballot(mine)) <= 2) break;
        unsigned long long v = mine ? q : 0ull;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((int)__lane_id() == leader) atomicAdd(pend + g, v);
        if (mine) add = false;
        m = __ballot(add);
    }
    if (add) atomicAdd(pend + gidx, q);
}
