// One path vertex of the wavefront path tracer, shared by the single-job
// bounce loop (mt_render) and the multi-job wavefront engine (mt_batch), so
// that a path computes the same bits whichever driver runs it.
//
// Restates CDeviceGI.loop's per-bounce kernels (src/cl/device/
// CDeviceGI.java:214-271): IntersectPrimitives' hit shading (NormalBVH.cl:
// 124-220), SetupBSDFPath (Render.cl:10-40), LightHitPass (:42-125),
// DirectLight (:377-466) via illuminateLight (:317-375) and
// SampleBSDFRayDirection (:233-285), plus the camera rays of
// InitCameraRayData (Raytrace.cl:1-32).
//
// Differences from the reference, each on purpose:
//  - RNG seeds are WangHash inputs keyed by (job seed, pid, bounce,
//    purpose) instead of (thread id, host random primes) (Sampling.cl:32-40,
//    CDeviceGI.java:364-376), so a path's samples do not depend on its queue
//    slot or on other jobs;
//  - paths whose sampled BSDF value or pdf is zero end at once (the
//    reference keeps tracing them; their later contributions are zero or
//    filtered as NaN by addAccum, Render.cl:1-7), and Russian roulette is
//    optional (the reference always traces a fixed 15 bounces);
//  - addAccum's validity test is applied to the rgb lanes only;
//  - the camera is a pinhole with per-sample jitter inside the pixel
//    (InitCameraRayDataJitter, Initialize.cl:129-172, jitters by up to one
//    pixel in NDC through a camera matrix).
#pragma once
#include "mt_bsdf.h"
#include "mt_sat.h"

namespace {

struct Camera {
    float ox, oy, oz;
    float fx, fy, fz, rx, ry, rz, ux, uy, uz;
    float tan_half, aspect;
    int w, h, spp;
};

// Pinhole basis from a render description (host, double precision).
inline int32_t camera_from_desc(const mt_render_desc* desc, Camera& C)
{
    double f[3] = {(double)desc->look[0] - desc->pos[0], (double)desc->look[1] - desc->pos[1],
                   (double)desc->look[2] - desc->pos[2]};
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    MT_ARG_CHECK(fl > 0.0, "camera: look-at equals position");
    for (double& x : f) x /= fl;
    const float* up = desc->up;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    MT_ARG_CHECK(rl > 0.0, "camera: up vector parallel to the view direction");
    for (double& x : r) x /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    C.ox = desc->pos[0]; C.oy = desc->pos[1]; C.oz = desc->pos[2];
    C.fx = (float)f[0]; C.fy = (float)f[1]; C.fz = (float)f[2];
    C.rx = (float)r[0]; C.ry = (float)r[1]; C.rz = (float)r[2];
    C.ux = (float)u[0]; C.uy = (float)u[1]; C.uz = (float)u[2];
    C.tan_half = (float)tan(0.5 * desc->vfov_deg * M_PI / 180.0);
    C.aspect = (float)desc->width / (float)desc->height;
    C.w = desc->width;
    C.h = desc->height;
    C.spp = desc->spp;
    return MT_OK;
}

struct SceneView {
    const float4* points;
    const float4* normals;
    const int* faces;
    const MtMaterial* mats;
    const int* light_faces;
    const int* num_lights;
    EnvView env;
    LgView G; // light-grid guided environment light (G.mode 0: off)
};

// RNG stream of one path at one bounce: generate_seed's form
// id + key * prime (Sampling.cl:32-40) with id = path id.
__device__ __forceinline__ void seed_for(uint32_t seed, int pid, int bounce, int purpose, int32_t& x,
                                         int32_t& y)
{
    int32_t key = (int32_t)((seed * 64u + (uint32_t)bounce) * 8u + (uint32_t)purpose + 1u);
    x = (int32_t)((uint32_t)pid + (uint32_t)key * 1000000007u);
    y = (int32_t)((uint32_t)pid + (uint32_t)key * 998244353u);
}

// Path-id convention: pid = k * spp + sample, where k numbers the pixels
// in bands of 8 rows, column by column inside a band (the last band holds
// h % 8 rows).  Consecutive ids are the samples of one pixel, then the
// pixels of 8x8 tiles, so the camera pass (slot = pid) traces neighbouring
// rays in every wave and the resolve reads each pixel's samples
// contiguously.  A path's result depends only on its id, never on its slot.
__host__ __device__ inline void path_pixel(int pid, int w, int h, int spp, int& x, int& y)
{
    const int k = pid / spp;
    const int full = (h >> 3) * 8 * w; // pixels in full bands
    int band, rows, r;
    if (k < full) {
        band = k / (8 * w);
        rows = 8;
        r = k - band * 8 * w;
    } else {
        band = h >> 3;
        rows = h - 8 * band;
        r = k - full;
    }
    x = r / rows;
    y = band * 8 + (r - x * rows);
}

// First path id of pixel (x, y) (its samples follow it).
__host__ __device__ inline int pixel_first_path(int x, int y, int w, int h, int spp)
{
    const int band = y >> 3, full_bands = h >> 3;
    const int k = band < full_bands ? band * 8 * w + x * 8 + (y & 7)
                                    : full_bands * 8 * w + x * (h - 8 * full_bands) + (y - 8 * band);
    return k * spp;
}

// Primary ray of path `pid` (pixel from path_pixel), jittered inside its
// pixel; tMin 0.001 and tMax inf as initGlobalRay (Util.cl:577-587).
__device__ __forceinline__ void camera_ray(const Camera& C, int pid, uint32_t seed, float4& o_tmin,
                                           float4& d_tmax)
{
    int px, py;
    path_pixel(pid, C.w, C.h, C.spp, px, py);
    int32_t sx, sy;
    seed_for(seed, pid, 0, 0, sx, sy);
    float j0 = mt_get_random(sx, sy);
    float j1 = mt_get_random(sx, sy);
    float a = (2.0f * ((float)px + j0) / (float)C.w - 1.0f) * C.tan_half * C.aspect;
    float b = (1.0f - 2.0f * ((float)py + j1) / (float)C.h) * C.tan_half;
    v3 d = norm3(mk3(C.fx + a * C.rx + b * C.ux, C.fy + a * C.ry + b * C.uy, C.fz + a * C.rz + b * C.uz));
    o_tmin = make_float4(C.ox, C.oy, C.oz, 0.001f);
    d_tmax = make_float4(d.x, d.y, d.z, INFINITY);
}

__device__ __forceinline__ float mis2(float a, float b) { return a / (a + b); } // Sampling.cl:206-215

__device__ __forceinline__ bool rgb_valid(v3 v)
{ // isFloat4AbsValid on rgb, Util.cl:537-549
    return !(isnan(v.x) || isinf(v.x) || v.x < 0.0f || isnan(v.y) || isinf(v.y) || v.y < 0.0f ||
             isnan(v.z) || isinf(v.z) || v.z < 0.0f);
}

struct AreaLight { // Light.cl:5-13
    v3 p1, e1, e2;
    MtFrame F;
    v3 intensity;
    float inv_area;
};

__device__ __forceinline__ AreaLight get_area_light(const SceneView& V, int tri)
{ // getAreaLight, Light.cl:43-77
    AreaLight A;
    const int* F = V.faces + 10 * tri;
    v3 p1 = xyz(V.points[F[0]]), p2 = xyz(V.points[F[1]]), p3 = xyz(V.points[F[2]]);
    v3 n;
    if (F[6] > -1 && V.normals) {
        v3 s = add3(add3(xyz(V.normals[F[6]]), xyz(V.normals[F[7]])), xyz(V.normals[F[8]]));
        n = norm3(mk3(s.x / 3.0f, s.y / 3.0f, s.z / 3.0f));
    } else {
        n = norm3(cross3(sub3(p2, p1), sub3(p3, p1)));
    }
    v3 e1 = sub3(p2, p1), e2 = sub3(p3, p1);
    A.inv_area = 2.0f / len3(cross3(e1, e2));
    A.F = mt_get_frame(n);
    A.p1 = p1;
    A.e1 = e1;
    A.e2 = e2;
    const MtMaterial& m = V.mats[F[9] & 0xFFFF];
    float pw = m.emission_param.y;
    A.intensity = mk3(m.emission_color.x * pw, m.emission_color.y * pw, m.emission_color.z * pw);
    return A;
}

__device__ __forceinline__ bool is_emitter(const MtMaterial& m)
{ // isBsdfEmitter / isSurfaceEmitter, Bsdf.cl:162-166, Material.cl:27-31
    float s = m.emission_param.x;
    return luminance3(m.emission_color.x * s, m.emission_color.y * s, m.emission_color.z * s) != 0.0f;
}

// What one bounce of one path produced.
struct PathOut {
    bool alive;          // the path continues with ro/rd/T
    float4 T;            // throughput.xyz, BSDF pdf of the sampled direction
    float4 ro, rd;       // next extension ray
    v3 Ladd;             // radiance gathered at this vertex (emitter / environment hit)
    bool has_shadow;     // a shadow ray so/sd carries the unoccluded contribution
    float4 so, sd, contrib;
    // guided environment mode: light-grid accumulation of this vertex
    int acc_gidx;                 // escaped ray: entry of (cell, tile), -1 none
    unsigned long long acc_q;     //   and its fixed-point luminance
    int shadow_gidx;              // environment light sample: entry the shadow ray adds
    float shadow_lum;             //   this luminance to when unoccluded (-1 none)
};

// LightHitPass, environment branch (Render.cl:106-123): the radiance a path
// that escaped the scene gathers (Ladd), and in the guided environment mode
// its light-grid entry and fixed-point luminance (acc_gidx, acc_q; -1 / 0
// otherwise).
__device__ __forceinline__ void shade_escape(const SceneView& V, int bounce, float4 ro, float4 rd, float4 T4,
                                             v3& Ladd, int& acc_gidx, unsigned long long& acc_q)
{
    const v3 d = xyz(rd);
    const v3 T = xyz(T4);
    const float last_pdf = T4.w;
    const bool last_specular = bounce == 0; // InitPathData sets it, SampleBSDFRayDirection clears it
    Ladd = mk3(0, 0, 0);
    acc_gidx = -1;
    acc_q = 0ull;
    if (V.env.w > 0) {
        float pdfw;
        float4 cv;
        if (V.G.mode) {
            // getRadianceEnvironmentLight1 (Light.cl:229-268) at the path's
            // last hit point: InitPathData's (0,0,0) for a camera ray
            // (Initialize.cl:46), else the ray's origin (SampleBSDFRayDirection
            // stores isect->p as hitpoint, Render.cl:278-279)
            const v3 hp = bounce == 0 ? mk3(0.0f, 0.0f, 0.0f) : xyz(ro);
            float lum;
            int gidx, cell, tile;
            cv = env_radiance_adaptive(V.env, mk3(V.G.cx, V.G.cy, V.G.cz), d, hp, pdfw, lum, gidx, cell, tile);
            if (V.G.mode == 2) pdfw = pdfw * lg_tile_pdf(V.G, cell, tile % LG_RX, tile / LG_RX);
            acc_gidx = gidx; // accumLightGrid, Light.cl:264
            acc_q = lg_fixed(lum);
        } else {
            int tx;
            cv = env_radiance(V.env, d, pdfw, tx);
        }
        v3 contrib = xyz(cv);
        if (!(contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f)) {
            float w = last_specular ? 1.0f : mis2(last_pdf, pdfw);
            v3 v = mk3(T.x * contrib.x * w, T.y * contrib.y * w, T.z * contrib.z * w);
            if (rgb_valid(v)) Ladd = v;
        }
    }
}

// One bounce of one path: the ray (ro, rd) of path `pid` at depth `bounce`
// with throughput T4 (w = last BSDF pdf) hit primitive c at distance tt
// (c < 0: miss).
__device__ __forceinline__ void shade_path(const SceneView& V, uint32_t seed, int rr_start, int flags, int pid,
                                           int bounce, float4 ro, float4 rd, float4 T4, int c, float tt,
                                           PathOut& out)
{
    v3 d = xyz(rd);
    v3 T = xyz(T4);
    float last_pdf = T4.w;
    bool last_specular = bounce == 0; // InitPathData sets it, SampleBSDFRayDirection clears it
    v3 Ladd = mk3(0, 0, 0);
    bool alive = false;
    out.has_shadow = false;
    out.so = make_float4(0, 0, 0, INFINITY);
    out.sd = make_float4(1, 0, 0, -1.0f);
    out.contrib = make_float4(0, 0, 0, 0);
    out.acc_gidx = -1;
    out.acc_q = 0ull;
    out.shadow_gidx = -1;
    out.shadow_lum = 0.0f;
    if (c < 0) {
        shade_escape(V, bounce, ro, rd, T4, Ladd, out.acc_gidx, out.acc_q);
    } else {
        // IntersectPrimitives hit shading (NormalBVH.cl:124-220)
        const int* F = V.faces + 10 * c;
        v3 p1 = xyz(V.points[F[0]]), p2 = xyz(V.points[F[1]]), p3 = xyz(V.points[F[2]]);
        v3 p = mk3(ro.x + tt * d.x, ro.y + tt * d.y, ro.z + tt * d.z);
        v3 e1 = sub3(p2, p1), e2 = sub3(p3, p1), e = sub3(p, p1);
        float d00 = dot3(e1, e1), d01 = dot3(e1, e2), d11 = dot3(e2, e2);
        float d20 = dot3(e, e1), d21 = dot3(e, e2);
        float den = (d00 * d11 - d01 * d01);
        float u = 0.0f, v = 0.0f;
        if (den != 0.0f) {
            float inv = 1.0f / den;
            u = (d11 * d20 - d01 * d21) * inv;
            v = (d00 * d21 - d01 * d20) * inv;
        }
        v3 n;
        if (F[6] > -1 && V.normals) {
            v3 n1 = xyz(V.normals[F[6]]), n2 = xyz(V.normals[F[7]]), n3 = xyz(V.normals[F[8]]);
            float w0 = (1.0f - u) - v;
            n = add3(add3(scl3(n1, w0), scl3(n2, u)), scl3(n3, v));
            if (dot3(n, d) > 0.00001f) {
                n = neg3(n);
                n1 = neg3(n1);
                n2 = neg3(n2);
                n3 = neg3(n3);
            }
            n = norm3(n);
            v3 tu = sub3(p, p1), tv = sub3(p, p2), tw = sub3(p, p3);
            v3 nn1 = norm3(n1), nn2 = norm3(n2), nn3 = norm3(n3);
            tu = sub3(tu, scl3(nn1, fminf(0.0f, dot3(tu, nn1))));
            tv = sub3(tv, scl3(nn2, fminf(0.0f, dot3(tv, nn2))));
            tw = sub3(tw, scl3(nn3, fminf(0.0f, dot3(tw, nn3))));
            p = add3(add3(add3(p, scl3(tu, w0)), scl3(tv, u)), scl3(tw, v));
        } else {
            n = norm3(cross3(e1, e2));
            if (dot3(n, d) > 0.00001f) n = neg3(n);
        }
        v3 ng = norm3(cross3(e1, e2));
        if (dot3(ng, d) > 0.00001f) ng = neg3(ng);
        int mat = F[9] & 0xFFFF;
        // SetupBSDFPath (Render.cl:10-40)
        MtBsdf B = mt_bsdf_setup_one(d, n, ng, mat, V.mats);
        if (B.mat >= 0) {
            const MtMaterial& M = V.mats[mat];
            int nlights = V.num_lights ? *V.num_lights : 0;
            if (is_emitter(M)) {
                // LightHitPass, emitter branch (Render.cl:76-104): the path ends here
                AreaLight AL = get_area_light(V, c);
                float cos_out = fmaxf(0.0f, dot3(AL.F.mz, neg3(d)));
                if (cos_out != 0.0f &&
                    !(AL.intensity.x == 0.0f && AL.intensity.y == 0.0f && AL.intensity.z == 0.0f)) {
                    float w = 1.0f;
                    if (!last_specular) {
                        float pick = 1.0f / (float)nlights;
                        float pdfw = (AL.inv_area * (tt * tt)) / fabsf(B.fix.z); // pdfAtoW, Sampling.cl:229-235
                        w = mis2(last_pdf, pdfw * pick);
                    }
                    v3 v = mk3(T.x * AL.intensity.x * w, T.y * AL.intensity.y * w, T.z * AL.intensity.z * w);
                    if (rgb_valid(v)) Ladd = v;
                }
            } else {
                // DirectLight (Render.cl:377-466) via illuminateLight (:317-375)
                int32_t sx, sy;
                seed_for(seed, pid, bounce, 1, sx, sy);
                // random_float4 (Sampling.cl:80-86): xy pick a light point or,
                // in the guided mode, the grid tile; zw the texel inside it
                float s0 = mt_get_random(sx, sy), s1 = mt_get_random(sx, sy);
                float s2 = mt_get_random(sx, sy), s3 = mt_get_random(sx, sy);
                int env_gidx = -1;
                float env_lum = 0.0f;
                float pick = 1.0f;
                bool use_env = V.env.w > 0, use_mesh = false;
                if (use_env && nlights > 0) {
                    float r = mt_get_random(sx, sy);
                    pick *= 0.5f;
                    use_mesh = r < 0.5f;
                    use_env = !use_mesh;
                } else if (!use_env && nlights > 0) {
                    use_mesh = true;
                }
                v3 ldir = mk3(0, 0, 0);
                float dist = 0.0f, lpdf = 0.0f;
                v3 rad = mk3(0, 0, 0);
                if (use_env && V.G.mode) { // illuminateEnvironmentLight1, Light.cl:140-193
                    float tpdf;
                    float4 cv = env_illuminate_adaptive(V.env, V.G.func, V.G.sat, V.G.tab,
                                                        mk3(V.G.cx, V.G.cy, V.G.cz), p,
                                                        make_float4(s0, s1, s2, s3), ldir, lpdf, env_lum, env_gidx,
                                                        tpdf);
                    if (V.G.mode == 2) lpdf = (tpdf > 0.0f && tpdf < INFINITY) ? lpdf * tpdf : 0.0f;
                    rad = xyz(cv);
                    dist = MT_FLOATMAX;
                } else if (use_env) { // illuminateEnvironmentLight, Light.cl:195-227
                    int tx;
                    float4 cv = env_illuminate(V.env, s0, s1, ldir, lpdf, tx);
                    rad = xyz(cv);
                    dist = MT_FLOATMAX;
                } else if (use_mesh) { // illuminateMeshLight, Render.cl:287-315
                    int li = (int)(mt_get_random(sx, sy) * (float)nlights);
                    int tri = (flags & MT_RENDER_PICKED_LIGHT) ? V.light_faces[li] : V.light_faces[0];
                    pick *= 1.0f / (float)nlights;
                    AreaLight AL = get_area_light(V, tri);
                    float su = sqrtf(s0); // sample_barycentric, Sampling.cl:182-186
                    float bu = 1.0f - su, bv = s1 * su;
                    v3 lp = add3(add3(AL.p1, scl3(AL.e1, bu)), scl3(AL.e2, bv));
                    v3 dv = sub3(lp, p);
                    float dist2 = dot3(dv, dv);
                    dist = sqrtf(dist2);
                    ldir = mk3(dv.x / dist, dv.y / dist, dv.z / dist);
                    float cn = dot3(AL.F.mz, neg3(ldir));
                    if (!(cn < MT_EPS_COSINE)) { // illuminateAreaLight, Light.cl:85-109
                        lpdf = (AL.inv_area * dist2) / cn;
                        rad = AL.intensity;
                    }
                }
                if (!(rad.x == 0.0f && rad.y == 0.0f && rad.z == 0.0f)) {
                    float bpdf = 0.0f, cosg = 0.0f;
                    v3 f = mt_bsdf_eval_one(B, M, ldir, bpdf, cosg);
                    if (!(f.x == 0.0f && f.y == 0.0f && f.z == 0.0f)) {
                        float w = mis2(lpdf, bpdf);
                        float s = (w * cosg) / (pick * lpdf);
                        v3 cv = mk3(s * (rad.x * f.x) * T.x, s * (rad.y * f.y) * T.y, s * (rad.z * f.z) * T.z);
                        if (rgb_valid(cv)) {
                            // initGlobalRayT(p, dir, dist - 2*EPS_RAY), Render.cl:453
                            out.so = make_float4(p.x, p.y, p.z, 0.0001f);
                            out.sd = make_float4(ldir.x, ldir.y, ldir.z, dist - 2 * MT_EPS_RAY);
                            out.contrib = make_float4(cv.x, cv.y, cv.z, 1.0f);
                            out.has_shadow = true;
                            // the commented-out accumLightGrid after an
                            // unoccluded test (Render.cl:458-459)
                            out.shadow_gidx = env_gidx;
                            out.shadow_lum = env_lum;
                        }
                    }
                }
                // SampleBSDFRayDirection (Render.cl:233-285)
                int32_t bx, by;
                seed_for(seed, pid, bounce, 2, bx, by);
                float u0 = mt_get_random(bx, by), u1 = mt_get_random(bx, by), u2 = mt_get_random(bx, by);
                v3 wdir;
                float pdf = 0.0f, cosg = 0.0f;
                v3 f = mt_bsdf_sample_one(B, M, u0, u1, u2, wdir, pdf, cosg);
                float k = cosg / pdf;
                v3 Tn = mk3(T.x * (f.x * k), T.y * (f.y * k), T.z * (f.z * k));
                alive = pdf > 0.0f && !(f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) && rgb_valid(Tn) &&
                        !(Tn.x == 0.0f && Tn.y == 0.0f && Tn.z == 0.0f);
                if (alive && rr_start >= 0 && bounce + 1 >= rr_start) {
                    float q = fminf(1.0f, fmaxf(Tn.x, fmaxf(Tn.y, Tn.z)));
                    float r = mt_get_random(bx, by);
                    if (r >= q) alive = false;
                    else Tn = mk3(Tn.x / q, Tn.y / q, Tn.z / q);
                }
                if (alive) {
                    out.T = make_float4(Tn.x, Tn.y, Tn.z, pdf);
                    out.ro = make_float4(p.x, p.y, p.z, 0.001f); // initGlobalRay
                    out.rd = make_float4(wdir.x, wdir.y, wdir.z, INFINITY);
                }
            }
        }
    }
    out.alive = alive;
    out.Ladd = Ladd;
}

// Position of this lane's entry when the wave's flagged lanes append to a
// dense queue: one atomic per wave, lane order inside the wave.  Call with
// the whole wave active.  Queue order across waves follows arrival, which
// no result depends on (each path owns at most one entry per bounce).
__device__ __forceinline__ int wave_append(bool flag, int* counter)
{
    const uint64_t m = __ballot(flag);
    int base = 0;
    if (m) {
        if (__lane_id() == 0) base = atomicAdd(counter, (int)__popcll(m));
        base = __shfl(base, 0);
    }
    return base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Shadow-ray and path queues are segmented (see MtSegQueue): work-group b of
// a launch appends to segment b % NSEG, whose capacity is the slots of the
// work-groups that map to it.  The render loop uses SHADOW_SEGS segments
// (about one resident producer group per segment); the multi-job engine,
// whose per-iteration batches are smaller, its own count.
constexpr int SHADOW_SEGS = 1024;
// Work-groups per segment in the fused loop's launches that sweep a
// segmented queue: group b takes segment b % SHADOW_SEGS and walks its
// chunks b / SHADOW_SEGS, + SEG_SWEEP, ...  The grid (SHADOW_SEGS x
// SEG_SWEEP) does not scale with the queue's capacity, so a nearly empty
// late bounce does not dispatch a capacity's worth of empty groups (about
// 27 us per launch at 130K groups).  A group appends only what it read from
// its own segment, which keeps every segment within its capacity.
constexpr int SEG_SWEEP = 32;

template <int NSEG = SHADOW_SEGS>
inline int shadow_seg_cap(int64_t slots, int group)
{
    const int64_t groups = (slots + group - 1) / group;
    return (int)((groups + NSEG - 1) / NSEG) * group;
}

// Entry of this lane in a segmented queue walked chunk-major: work-group b
// takes chunk b / SHADOW_SEGS of segment b % SHADOW_SEGS (seg_cap is a
// multiple of the group size), so the live chunks at the front of every
// segment are dispatched first.
template <int NSEG = SHADOW_SEGS>
__device__ __forceinline__ int seg_entry(int seg_cap)
{
    const int p = blockIdx.x % NSEG, c = blockIdx.x / NSEG;
    return p * seg_cap + c * blockDim.x + threadIdx.x;
}

// Whether entry `id` of a segmented queue of NSEG segments holds an item.
template <int NSEG = SHADOW_SEGS>
__device__ __forceinline__ bool seg_live(const int* __restrict__ counts, int seg_cap, int id)
{
    const int p = id / seg_cap;
    return p < NSEG && id - p * seg_cap < counts[p];
}

// Per-pixel mean over samples in sample order (averageAccum, Image.cl:29-45,
// without tone mapping): deterministic, no float atomics.  A pixel's samples
// are consecutive path ids, so a work-group takes a run of consecutive pixel
// ranks, stages their paths' radiance through LDS with coalesced loads, and
// thread t then sums pixel t's samples in order (the same sums as reading
// them one by one).
constexpr int RESOLVE_PATHS = 1024; // radiance values staged per round
__host__ __device__ inline int resolve_pixels_per_group(int spp)
{
    const int p = spp >= RESOLVE_PATHS ? 1 : RESOLVE_PATHS / spp;
    return p < 256 ? p : 256;
}

__global__ __launch_bounds__(256) void k_resolve(const float4* __restrict__ L, int w, int h, int spp,
                                                 float4* __restrict__ image)
{
    __shared__ float4 s_L[RESOLVE_PATHS];
    const int npix = w * h;
    const int ppg = resolve_pixels_per_group(spp);
    const int k0 = blockIdx.x * ppg;
    if (k0 >= npix) return;
    const int npx = npix - k0 < ppg ? npix - k0 : ppg;
    const int t = threadIdx.x;
    const float4* Lg = L + (size_t)k0 * spp;
    const int total = npx * spp; // > RESOLVE_PATHS only when one pixel has more samples
    float r = 0.0f, g = 0.0f, b = 0.0f;
    for (int c = 0; c < total; c += RESOLVE_PATHS) {
        const int m = total - c < RESOLVE_PATHS ? total - c : RESOLVE_PATHS;
        __syncthreads();
        for (int i = t; i < m; i += blockDim.x) s_L[i] = Lg[c + i];
        __syncthreads();
        if (t < npx) {
            int lo = t * spp, hi = lo + spp;
            lo = lo > c ? lo : c;
            hi = hi < c + m ? hi : c + m;
            for (int i = lo; i < hi; i++) {
                const float4 v = s_L[i - c];
                r += v.x;
                g += v.y;
                b += v.z;
            }
        }
    }
    if (t < npx) {
        int x, y;
        path_pixel((k0 + t) * spp, w, h, spp, x, y);
        float inv = 1.0f / (float)spp;
        image[(size_t)y * w + x] = make_float4(r * inv, g * inv, b * inv, 1.0f);
    }
}

} // namespace
