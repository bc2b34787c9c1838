/*
 * ORACLE (test infrastructure only) — stage C: light-grid SAT build and
 * sampling, environment-map SAT and environment-light sampling.
 *
 * Parity status
 *   orc_sat_tiles        restates prefixSumRowSubgrid / prefixSumColSubgrid
 *                        (LightGrid.cl:347-400; 16x16 copy test/lightgrid/
 *                        SimpleGrid.cl:16-69): inclusive Kogge-Stone scans of
 *                        rx consecutive entries, then of ry entries down each
 *                        column segment, in the kernels' exact summation tree.
 *                        Pinned by the SimpleGrid impulse known answer
 *                        (TestSimpleGrid.java:39-56) and the all-ones
 *                        (x+1)(y+1) known answer (TestLightGrid.java:41-62),
 *                        plus a float64 brute-force SAT on random tiles.
 *   orc_lg_sample        restates sampleSubgridContinuous and helpers
 *                        (LightGrid.cl:39-320).  Pinned against a pbrt
 *                        piecewise-constant Distribution2D on the same tile
 *                        (test/sampling/Distribution2D.java) computed in
 *                        float64 by the tests: equal cell picks away from CDF
 *                        boundaries, uv and pdf within 1e-4.
 *   orc_lg_set_tile_func restates CLightGrid.setToAllFunc (CLightGrid.java:
 *                        86-100); the local index -> (x, y) mapping of the
 *                        un-vendored SATSubgrid is our choice (x = i % 16,
 *                        y = i / 16): parity unpinned.
 *   orc_env_sat          restates calculateSAT (SAT.cl:304-310): sequential
 *                        row scans, then sequential column scans.  The host
 *                        SAT class actually used (CEnvironment.java:48-55) is
 *                        un-vendored: parity unpinned beyond this kernel order;
 *                        pinned by a float64 SAT (closed form) within 1e-6 rel.
 *   orc_sat_sample_region, orc_env_illuminate, orc_env_radiance restate
 *                        SAT.cl:95-342 and Light.cl:129-137,195-227,270-286;
 *                        pinned against Distribution2D (float64) and the
 *                        closed-form uniform-map case (pdf == 1 everywhere).
 *   orc_env_illuminate_adaptive, orc_env_radiance_adaptive restate the
 *                        light-grid-guided environment light (Light.cl:35-38,
 *                        140-193, 229-268; no live caller in the reference).
 *                        Pinned by structure: the accum entry names the
 *                        camera cell and ray tile of an independent float64
 *                        spherical binning; the tile equals orc_lg_sample's
 *                        pick; the texel lies in the tile's closed-form
 *                        rectangle and equals region sampling of it; a
 *                        constant map gives pdf 1 / (2 pi^2 sin theta) and
 *                        accum = count * luminance.
 *
 * Deviations (documented): getSphericalGridXY (Util.cl:397-410) yields
 * y == height for d.y == -1 and getFunc/getCummulative index -1 for columns
 * past the map (SAT.cl:81-105); the reference then reads outside its buffers.
 * The oracle clamps the texel index to the map and returns 0 for such func
 * reads.  getEnvironmentLight (Light.cl:129-137) calls setRegion before the
 * region's nu/nv are set; the intended full-map region (0,0,w,h) is used.
 */
#include "orc.h"
#include "orc_math.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Kogge-Stone tile SAT: LightGrid.cl:347-400                          */
/* ------------------------------------------------------------------ */

static void kogge_stone(float* a, int n)
{ /* aux[l] = aux[l-o] + aux[l] for l >= o, o = 1,2,4,... (reads before writes) */
    for (int o = 1; o < n; o *= 2)
        for (int l = n - 1; l >= o; l--) a[l] = a[l - o] + a[l];
}

int orc_sat_tiles(const float* func, float* sat, int32_t nu, int32_t nv, int32_t rx, int32_t ry)
{
    if (nu <= 0 || nv <= 0 || rx <= 0 || ry <= 0 || ry > 1024 || rx > 1024) return ORC_ERR_ARG;
    int64_t total = (int64_t)nu * nv;
    if (total % rx || total % ry || nv % ry) return ORC_ERR_ARG;
    float aux[1024];
    /* prefixSumRowSubgrid: work-group g covers entries g*rx .. g*rx+rx-1 */
    for (int64_t g = 0; g < total / rx; g++) {
        for (int l = 0; l < rx; l++) aux[l] = func[g * rx + l];
        kogge_stone(aux, rx);
        for (int l = 0; l < rx; l++) sat[g * rx + l] = aux[l];
    }
    /* prefixSumColSubgrid: i = ry*g, column i / nv, rows i % nv + l */
    for (int64_t g = 0; g < total / ry; g++) {
        int64_t i = (int64_t)ry * g;
        int64_t xi = i / nv;
        int64_t yi0 = i % nv;
        for (int l = 0; l < ry; l++) aux[l] = sat[xi + (yi0 + l) * nu];
        kogge_stone(aux, ry);
        for (int l = 0; l < ry; l++) sat[xi + (yi0 + l) * nu] = aux[l];
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Light grid index math and sampling: LightGrid.cl:39-320             */
/* ------------------------------------------------------------------ */

enum { LG_NU = 1600, LG_NV = 1600, LG_RX = 16, LG_RY = 32, LG_CX = 100, LG_CY = 50 };

static int lg_global_index(int s, int lx, int ly)
{ /* globalIndexInSubgrid, LightGrid.cl:104-139 */
    int cx = s % LG_CX;                 /* subgridGlobalIndex1 */
    int cy = s / LG_CX;
    int gi = cx * LG_RX + cy * (LG_RX * LG_RY) * LG_CX; /* subgridGlobalIndex2 */
    int gx = gi % LG_NU;                /* subgridGlobalIndexXY */
    int gy = gi / LG_NV;
    return (gx + lx) + (gy + ly) * LG_NU;
}

static float lg_sat(const float* sat, int s, int sx, int sy)
{ /* getSubgridValueSAT, :142-148 (out-of-range reads return 0) */
    int in_range = !(sx >= LG_RX || sx < 0) && !(sy >= LG_RY || sy < 0);
    return in_range ? sat[lg_global_index(s, sx, sy)] : 0.0f;
}

static float lg_range(const float* sat, int s, int minx, int miny, int maxx, int maxy)
{ /* getSubgridSATRange, :150-157 */
    float a1 = lg_sat(sat, s, maxx, maxy);
    float b1 = lg_sat(sat, s, minx - 1, maxy);
    float c1 = lg_sat(sat, s, maxx, miny - 1);
    float d1 = lg_sat(sat, s, minx - 1, miny - 1);
    return ((a1 + d1) - b1) - c1;
}

static float lg_func_int(const float* sat, int s, int y) { return lg_range(sat, s, 0, y, LG_RX - 1, y); }

static float lg_cond(const float* sat, int s, int x, int y)
{ /* getSubgridConditional, :183-187 */
    float fi = lg_func_int(sat, s, y);
    return lg_range(sat, s, 0, y, x - 1, y) / fi;
}

static float lg_marg(const float* sat, int s, int y)
{ /* getSubgridMarginal, :189-193 */
    float last = lg_sat(sat, s, LG_RX - 1, LG_RY - 1);
    return lg_sat(sat, s, LG_RX - 1, y - 1) / last;
}

static int lg_upper_cond(const float* sat, int s, int y, int first, int last, float v)
{ /* upperSubgridBoundConditional, :212-229 */
    int b = first, e = last;
    while (b != e) {
        int mid = b + (e - b) / 2;
        float mv = lg_cond(sat, s, mid, y);
        int right = !(v < mv);
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    return b;
}

static int lg_upper_marg(const float* sat, int s, int first, int last, float v)
{ /* upperSubgridBoundMarginal, :233-247 */
    int b = first, e = last;
    while (b != e) {
        int mid = b + (e - b) / 2;
        float mv = lg_marg(sat, s, mid);
        int right = !(v < mv);
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    return b;
}

static void lg_sample_one(const float* func, const float* sat, int s, float u0, float u1, int* ox_out,
                          int* oy_out, float* uu_out, float* vv_out, float* pdf_out)
{ /* sampleSubgridContinuous, :310-320 */
    /* sampleSubgridContinuousMarginal, :279-296 */
    int ptr = lg_upper_marg(sat, s, 0, LG_RY, u1);
    int oy = ptr - 1 > 0 ? ptr - 1 : 0;
    float my0 = lg_marg(sat, s, oy);
    float du = (u1 - my0) / (lg_marg(sat, s, oy + 1) - my0);
    float last = lg_sat(sat, s, LG_RX - 1, LG_RY - 1);
    float pdf_v = (lg_func_int(sat, s, oy) * (float)(LG_RX * LG_RY)) / last;
    float v = ((float)oy + du) / (float)LG_RY;
    /* sampleSubgridContinuousConditional, :249-266 */
    ptr = lg_upper_cond(sat, s, oy, 0, LG_RX, u0);
    int ox = ptr - 1 > 0 ? ptr - 1 : 0;
    float cx0 = lg_cond(sat, s, ox, oy);
    float du2 = (u0 - cx0) / (lg_cond(sat, s, ox + 1, oy) - cx0);
    float pdf_u = func[lg_global_index(s, ox, oy)] / lg_func_int(sat, s, oy);
    *ox_out = ox;
    *oy_out = oy;
    *uu_out = ((float)ox + du2) / (float)LG_RX;
    *vv_out = v;
    *pdf_out = pdf_u * pdf_v;
}

int orc_lg_sample(const float* func, const float* sat, int32_t n, const int32_t* cell,
                  const float* u, int32_t* tile_xy, float* uv, float* pdf)
{
    for (int32_t i = 0; i < n; i++) {
        int s = cell[i];
        if (s < 0 || s >= LG_CX * LG_CY) return ORC_ERR_ARG;
        int ox, oy;
        lg_sample_one(func, sat, s, u[2 * i], u[2 * i + 1], &ox, &oy, &uv[2 * i], &uv[2 * i + 1], &pdf[i]);
        tile_xy[2 * i] = ox;
        tile_xy[2 * i + 1] = oy;
    }
    return ORC_OK;
}

int orc_lg_set_tile_func(float* func, const float* tile)
{ /* CLightGrid.setToAllFunc, CLightGrid.java:86-100 */
    for (int s = 0; s < LG_CX * LG_CY; s++)
        for (int i = 0; i < LG_RX * LG_RY; i++)
            func[lg_global_index(s, i % LG_RX, i / LG_RX)] = tile[i];
    return ORC_OK;
}

static void spherical_xy(int w, int h, f4 d, int* x, int* y)
{ /* getSphericalGridXY, Util.cl:397-410 */
    float phi = mt_xm_acosf(d.y);
    float theta = mt_xm_atan2f(d.z, d.x);
    *x = (int)((0.5f - (0.5f * theta) / ORC_PI) * (float)w);
    *y = (int)((phi / ORC_PI) * (float)h);
}

static int spherical_index(int w, int h, f4 d)
{ /* getSphericalGridIndex, Util.cl:413-417 (clamped to the map) */
    int x, y;
    spherical_xy(w, h, d, &x, &y);
    int idx = (int)((float)x + (float)y * (float)w);
    if (idx < 0) idx = 0;
    if (idx > w * h - 1) idx = w * h - 1;
    return idx;
}

int orc_lg_cell_from_dirs(int32_t n, const float* dirs, int32_t* cell, int32_t* tile_xy)
{ /* subgridIndexFromCamera (:68-78) on normalize(dir) and tileIndexXYFromCamera (:80-90) */
    for (int32_t i = 0; i < n; i++) {
        f4 v = f4_make(dirs[4 * i], dirs[4 * i + 1], dirs[4 * i + 2], 0.0f);
        f4 d = normalize3(v);
        cell[i] = spherical_index(LG_CX, LG_CY, d);
        int t = spherical_index(LG_RX, LG_RY, v);
        tile_xy[2 * i] = t % LG_RX;
        tile_xy[2 * i + 1] = t / LG_RX;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Environment map: Initialize.cl:100-127, SAT.cl, Light.cl            */
/* ------------------------------------------------------------------ */

int orc_env_lum(const float* rgba, int32_t w, int32_t h, float* lum)
{ /* InitEnvironmentLum, Initialize.cl:100-115 */
    for (int64_t i = 0; i < (int64_t)w * h; i++)
        lum[i] = luminance3(rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2]);
    return ORC_OK;
}

int orc_env_sat(const float* lum, int32_t w, int32_t h, float* sat)
{ /* InitEnvironmentSAT -> calculateSAT, SAT.cl:114-133,304-310 */
    memcpy(sat, lum, sizeof(float) * (size_t)w * h);
    for (int y = 0; y < h; y++) {
        float val = 0.0f;
        for (int x = 0; x < w; x++) { val = val + sat[(size_t)y * w + x]; sat[(size_t)y * w + x] = val; }
    }
    for (int x = 0; x < w; x++) {
        float val = 0.0f;
        for (int y = 0; y < h; y++) { val = val + sat[(size_t)y * w + x]; sat[(size_t)y * w + x] = val; }
    }
    return ORC_OK;
}

typedef struct { int nu, nv, minx, miny, sx, sy; } oregion;

static float r_value(const oregion* R, const float* sat, int x, int y)
{ /* getValueSAT, SAT.cl:135-147 */
    if (x < 0 || y < 0) return 0.0f;
    if (x > R->minx + R->sx - 1) return 0.0f;
    if (y > R->miny + R->sy - 1) return 0.0f;
    if (x > R->nu - 1 || y > R->nv - 1) return 0.0f;
    return sat[(size_t)y * R->nu + x];
}

static float r_range(const oregion* R, const float* sat, int minx, int miny, int maxx, int maxy)
{ /* getSATRange, SAT.cl:149-156 */
    float d1 = r_value(R, sat, maxx, maxy);
    float c1 = r_value(R, sat, minx - 1, maxy);
    float b1 = r_value(R, sat, maxx, miny - 1);
    float a1 = r_value(R, sat, minx - 1, miny - 1);
    return ((a1 + d1) - b1) - c1;
}

static int r_lastx(const oregion* R) { return R->minx + R->sx - 1; }
static int r_lasty(const oregion* R) { return R->miny + R->sy - 1; }
static float r_func_int(const oregion* R, const float* sat, int y)
{ return r_range(R, sat, R->minx, y, r_lastx(R), y); }
static float r_cond(const oregion* R, const float* sat, int x, int y)
{ float fi = r_func_int(R, sat, y); return r_range(R, sat, R->minx, y, x - 1, y) / fi; }
static float r_marg(const oregion* R, const float* sat, int y)
{
    float last = r_range(R, sat, R->minx, R->miny, r_lastx(R), r_lasty(R));
    return r_range(R, sat, R->minx, R->miny, r_lastx(R), y - 1) / last;
}
static float r_func(const oregion* R, const float* func, int x, int y)
{ /* getFunc, SAT.cl:95-99 (index -1 outside the map reads 0 here) */
    if (x < 0 || y < 0 || x > R->nu - 1 || y > R->nv - 1) return 0.0f;
    return func[(size_t)y * R->nu + x];
}
static float r_pdf_cond(const oregion* R, const float* sat, const float* func, int x, int y)
{ return r_func(R, func, x, y) / r_func_int(R, sat, y); }
static float r_pdf_marg(const oregion* R, const float* sat, int y)
{
    float fi = r_func_int(R, sat, y);
    float last = r_range(R, sat, R->minx, R->miny, r_lastx(R), r_lasty(R));
    return (fi * (float)(R->sx * R->sy)) / last;
}

static void r_sample(const oregion* R, const float* sat, const float* func, float u0, float u1,
                     float* uv, int* off, float* pdf)
{ /* sampleContinuous, SAT.cl:244-323 */
    int b = R->miny, e = R->miny + R->sy;
    while (b != e) { /* upperBoundMarginal, :227-242 */
        int mid = b + (e - b) / 2;
        float mv = r_marg(R, sat, mid);
        int right = !(u1 < mv);
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    int oy = b - 1 > 0 ? b - 1 : 0;
    float m0 = r_marg(R, sat, oy);
    float du = (u1 - m0) / (r_marg(R, sat, oy + 1) - m0);
    float pdf_v = r_pdf_marg(R, sat, oy);
    uv[1] = ((float)oy + du) / (float)R->nv;
    b = R->minx; e = R->minx + R->sx;
    while (b != e) { /* upperBoundConditional, :206-223 */
        int mid = b + (e - b) / 2;
        float mv = r_cond(R, sat, mid, oy);
        int right = !(u0 < mv);
        b = right ? mid + 1 : b;
        e = right ? e : mid;
    }
    int ox = b - 1 > 0 ? b - 1 : 0;
    float c0 = r_cond(R, sat, ox, oy);
    float du2 = (u0 - c0) / (r_cond(R, sat, ox + 1, oy) - c0);
    float pdf_u = r_pdf_cond(R, sat, func, ox, oy);
    uv[0] = ((float)ox + du2) / (float)R->nu;
    off[0] = ox;
    off[1] = oy;
    *pdf = pdf_u * pdf_v;
}

int orc_sat_sample_region(const float* func, const float* sat, int32_t nu, int32_t nv,
                          int32_t minx, int32_t miny, int32_t sx, int32_t sy,
                          int32_t n, const float* u, float* uv, int32_t* offset, float* pdf)
{
    if (minx < 0 || miny < 0 || sx <= 0 || sy <= 0 || minx + sx > nu || miny + sy > nv) return ORC_ERR_ARG;
    oregion R = {nu, nv, minx, miny, sx, sy};
    for (int32_t i = 0; i < n; i++) {
        int off[2];
        r_sample(&R, sat, func, u[2 * i], u[2 * i + 1], uv + 2 * i, off, pdf + i);
        offset[2 * i] = off[0];
        offset[2 * i + 1] = off[1];
    }
    return ORC_OK;
}

static f4 spherical_direction(float u, float v)
{ /* getSphericalDirection, Util.cl:423-439 */
    float phi = ORC_PI * v;
    float theta = (u * 2.0f) * ORC_PI;
    return f4_make(-mt_xm_sinf(phi) * mt_xm_cosf(theta), mt_xm_cosf(phi), mt_xm_sinf(phi) * mt_xm_sinf(theta), 0.0f);
}

int orc_env_illuminate(const float* rgba, const float* lum, const float* sat,
                       int32_t w, int32_t h, int32_t n, const float* u,
                       float* dir, float* contrib, float* pdfw, int32_t* texel)
{ /* illuminateEnvironmentLight, Light.cl:195-227 */
    oregion R = {w, h, 0, 0, w, h};
    for (int32_t i = 0; i < n; i++) {
        float uv[2], pdf;
        int off[2];
        r_sample(&R, sat, lum, u[2 * i], u[2 * i + 1], uv, off, &pdf);
        f4 d = spherical_direction(uv[0], uv[1]);
        int idx = spherical_index(w, h, d);
        float s = mt_xm_sinf(uv[1] * ORC_PI);
        float pw = pdf / (((2.0f * ORC_PI) * ORC_PI) * mt_xm_sinf(uv[1] * ORC_PI));
        if (s == 0.0f) pw = 0.0f;
        dir[4 * i] = d.x; dir[4 * i + 1] = d.y; dir[4 * i + 2] = d.z; dir[4 * i + 3] = 0.0f;
        for (int k = 0; k < 4; k++) contrib[4 * i + k] = rgba[4 * (size_t)idx + k];
        pdfw[i] = pw;
        texel[i] = idx;
    }
    return ORC_OK;
}

int orc_env_radiance(const float* rgba, const float* lum, const float* sat,
                     int32_t w, int32_t h, int32_t n, const float* dir,
                     float* contrib, float* pdfw, int32_t* texel)
{ /* getRadianceEnvironmentLight, Light.cl:270-286 */
    oregion R = {w, h, 0, 0, w, h};
    for (int32_t i = 0; i < n; i++) {
        f4 d = f4_make(dir[4 * i], dir[4 * i + 1], dir[4 * i + 2], 0.0f);
        int x, y;
        spherical_xy(w, h, d, &x, &y);
        int idx = spherical_index(w, h, d);
        float p = r_pdf_cond(&R, sat, lum, x, y) * r_pdf_marg(&R, sat, y); /* getPdfSAT, SAT.cl:336-342 */
        float v = (float)y / (float)h;
        pdfw[i] = p / (((2.0f * ORC_PI) * ORC_PI) * mt_xm_sinf(v * ORC_PI));
        for (int k = 0; k < 4; k++) contrib[4 * i + k] = rgba[4 * (size_t)idx + k];
        texel[i] = idx;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Adaptive environment light: Light.cl:35-38,140-193,229-268          */
/* ------------------------------------------------------------------ */

static int cell_from_camera(const float* cam, f4 hit)
{ /* subgridIndexFromCamera, LightGrid.cl:68-78 */
    f4 c = f4_make(cam[0], cam[1], cam[2], 0.0f);
    return spherical_index(LG_CX, LG_CY, normalize3(f4_sub(f4_make(hit.x, hit.y, hit.z, 0.0f), c)));
}

static oregion tile_region(int w, int h, int tx, int ty)
{ /* getSubgridUnitBound (LightGrid.cl:158-167), getSubRegionFromUnitBound (SAT.cl:28-44) */
    float ux0 = (float)tx / (float)LG_RX, uy0 = (float)ty / (float)LG_RY;
    float ux1 = (float)(tx + 1) / (float)LG_RX, uy1 = (float)(ty + 1) / (float)LG_RY;
    int minx = (int)((float)w * ux0), miny = (int)((float)h * uy0);
    int maxx = (int)((float)w * ux1), maxy = (int)((float)h * uy1);
    oregion R = {w, h, minx, miny, maxx - minx, maxy - miny};
    return R;
}

/* illuminateEnvironmentLight1, Light.cl:140-193, one sample; *tile_pdf =
 * the tile pick's continuous pdf (pdfLL, unused by the reference). */
void orc_env_illum_adaptive_one(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                                const float* lg_func, const float* lg_sat_arr, const float* cam, const float* hit,
                                const float* q, float* dir, float* contrib, float* pdfw, float* lum_out,
                                int32_t* gidx, int32_t* texel_xy, float* tile_pdf)
{
    int s = cell_from_camera(cam, f4_make(hit[0], hit[1], hit[2], 0.0f));
    int lx, ly;
    float luu, lvv, lpdf;
    lg_sample_one(lg_func, lg_sat_arr, s, q[0], q[1], &lx, &ly, &luu, &lvv, &lpdf);
    oregion R = tile_region(w, h, lx, ly);
    float uv[2], pdf;
    int off[2];
    r_sample(&R, sat, lum, q[2], q[3], uv, off, &pdf);
    f4 d = spherical_direction(uv[0], uv[1]);
    int idx = spherical_index(w, h, d);
    float sn = mt_xm_sinf(uv[1] * ORC_PI);
    float pw = pdf / (((2.0f * ORC_PI) * ORC_PI) * mt_xm_sinf(uv[1] * ORC_PI));
    if (sn == 0.0f) pw = 0.0f;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z; dir[3] = 0.0f;
    for (int k = 0; k < 4; k++) contrib[k] = rgba[4 * (size_t)idx + k];
    *pdfw = pw;
    *lum_out = luminance3(rgba[4 * (size_t)idx], rgba[4 * (size_t)idx + 1], rgba[4 * (size_t)idx + 2]);
    *gidx = lg_global_index(s, lx, ly);
    if (texel_xy) {
        texel_xy[0] = off[0];
        texel_xy[1] = off[1];
    }
    if (tile_pdf) *tile_pdf = lpdf;
}

int orc_env_illuminate_adaptive(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                                const float* lg_func, const float* lg_sat_arr, const float* cam, int32_t n,
                                const float* hit, const float* u, float* dir, float* contrib, float* pdfw,
                                float* lum_out, int32_t* gidx, int32_t* texel_xy)
{ /* illuminateEnvironmentLight1, Light.cl:140-193 */
    if (w < LG_RX || h < LG_RY) return ORC_ERR_ARG;
    for (int32_t i = 0; i < n; i++)
        orc_env_illum_adaptive_one(rgba, lum, sat, w, h, lg_func, lg_sat_arr, cam, hit + 4 * (size_t)i,
                                   u + 4 * (size_t)i, dir + 4 * (size_t)i, contrib + 4 * (size_t)i, pdfw + i,
                                   lum_out + i, gidx + i, texel_xy ? texel_xy + 2 * (size_t)i : NULL, NULL);
    return ORC_OK;
}

/* getRadianceEnvironmentLight1, Light.cl:229-268, one escaped ray, without
 * the accumulation; *cell / *tile = the grid cell of the hit point and the
 * tile (ty * 16 + tx) of the direction. */
void orc_env_rad_adaptive_one(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                              const float* cam, const float* dir, const float* hit, float* contrib, float* pdfa,
                              float* lum_out, int32_t* gidx, int32_t* cell, int32_t* tile)
{
    f4 d = f4_make(dir[0], dir[1], dir[2], 0.0f);
    int s = cell_from_camera(cam, f4_make(hit[0], hit[1], hit[2], 0.0f));
    int t = spherical_index(LG_RX, LG_RY, d); /* tileIndexXYFromCamera, LightGrid.cl:80-90 */
    int tx = t % LG_RX, ty = t / LG_RX;
    oregion R = tile_region(w, h, tx, ty);
    int x, y;
    spherical_xy(w, h, d, &x, &y);
    float p = r_pdf_cond(&R, sat, lum, x, y) * r_pdf_marg(&R, sat, y); /* getPdfSAT, SAT.cl:336-342 */
    float sv = (float)y / (float)h;
    *pdfa = p / (((2.0f * ORC_PI) * ORC_PI) * mt_xm_sinf(sv * ORC_PI));
    int idx = spherical_index(w, h, d);
    for (int k = 0; k < 4; k++) contrib[k] = rgba[4 * (size_t)idx + k];
    *lum_out = luminance3(rgba[4 * (size_t)idx], rgba[4 * (size_t)idx + 1], rgba[4 * (size_t)idx + 2]);
    *gidx = lg_global_index(s, tx, ty);
    if (cell) *cell = s;
    if (tile) *tile = t;
}

int orc_env_radiance_adaptive(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                              float* accum, const float* cam, int32_t n, const float* dir, const float* hit,
                              float* contrib, float* pdfa, float* lum_out, int32_t* gidx)
{ /* getRadianceEnvironmentLight1, Light.cl:229-268; accumLightGrid :35-38 in index order */
    if (w < LG_RX || h < LG_RY) return ORC_ERR_ARG;
    for (int32_t i = 0; i < n; i++) {
        orc_env_rad_adaptive_one(rgba, lum, sat, w, h, cam, dir + 4 * (size_t)i, hit + 4 * (size_t)i,
                                 contrib + 4 * (size_t)i, pdfa + i, lum_out + i, gidx + i, NULL, NULL);
        if (accum) accum[gidx[i]] += lum_out[i];
    }
    return ORC_OK;
}

/* Continuous pdf of tile (tx, ty) in cell s, the value sampleSubgridContinuous
 * returns when it picks the tile: (func / rowInt) * (rowInt * 512 / total);
 * 0 for a tile the cell cannot pick. */
float orc_lg_tile_pdf(const float* lg_func, const float* lg_sat_arr, int32_t s, int32_t tx, int32_t ty)
{
    float fi = lg_func_int(lg_sat_arr, s, ty);
    float last = lg_sat(lg_sat_arr, s, LG_RX - 1, LG_RY - 1);
    float f = lg_func[lg_global_index(s, tx, ty)];
    float p = (f / fi) * ((fi * (float)(LG_RX * LG_RY)) / last);
    return (fi > 0.0f && last > 0.0f && p > 0.0f && p < INFINITY) ? p : 0.0f;
}

/* accumLightGrid in fixed point (units of 2^-20, round half up). */
uint64_t orc_lg_fixed(float lum)
{
    double v = (double)lum * 1048576.0;
    return (v >= 0.0 && v < 9.0e15) ? (uint64_t)floor(v + 0.5) : 0u;
}

/* End of a frame: pend folds into accum (one rounding per entry), cleared. */
int orc_lg_fold(uint64_t* pend, float* accum)
{
    for (int64_t i = 0; i < (int64_t)LG_NU * LG_NV; i++)
        if (pend[i]) {
            accum[i] = accum[i] + (float)((double)pend[i] * (1.0 / 1048576.0));
            pend[i] = 0;
        }
    return ORC_OK;
}
