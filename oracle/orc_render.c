/*
 * ORACLE (test infrastructure only) — the wavefront bounce loop run one path
 * at a time: CDeviceGI.loop (src/cl/device/CDeviceGI.java:214-271) with
 * SetupBSDFPath, LightHitPass, DirectLight and SampleBSDFRayDirection
 * (src/cl/kernel/Render.cl:10-125,233-466), built only from the oracle's
 * own stage functions (orc_trace_*, orc_shade_hits, orc_bsdf_*,
 * orc_env_*, orc_get_random) so that it checks the composed product loop.
 * It applies the same documented departures as the product loop (RNG keyed
 * by path id, early end of zero-throughput paths, optional Russian
 * roulette, pinhole camera with in-pixel jitter).
 *
 * The light-grid guided environment mode (lg_mode 1, 2) restates
 * illuminateEnvironmentLight1 / getRadianceEnvironmentLight1 with
 * accumLightGrid (Light.cl:35-38,140-268; the pairing Render.cl:458-459
 * leaves commented out) on the oracle's own grid arrays; grid sums are kept
 * in fixed point (2^-20 units) so they do not depend on thread order.
 *
 * Parity status: per-path results follow the product's arithmetic exactly
 * (the transcendental functions are the shared mt_xmath.h definitions), so
 * images compare bit for bit.  The loop itself is pinned by closed forms in
 * tests/test_oracle_render_pins.py: furnace (a Lambertian plane under a
 * constant environment, two segments, mean = the quadrature of the MIS
 * estimator, within 1% of albedo * radiance) for the full-map light and
 * both guided variants, the reference's guided pdf landing on its
 * closed-form bias; a triangle emitter's Lambert polygon irradiance for
 * illuminateMeshLight, pdfAtoW and the emitter MIS.
 */
#include "orc.h"
#include "orc_math.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static void seed_for(uint32_t seed, int32_t pid, int bounce, int purpose, int32_t st[2])
{
    int32_t key = (int32_t)((seed * 64u + (uint32_t)bounce) * 8u + (uint32_t)purpose + 1u);
    st[0] = (int32_t)((uint32_t)pid + (uint32_t)key * 1000000007u);
    st[1] = (int32_t)((uint32_t)pid + (uint32_t)key * 998244353u);
}

static float mis2(float a, float b) { return a / (a + b); }

static int rgb_valid(const float* v)
{
    for (int k = 0; k < 3; k++)
        if (isnan(v[k]) || isinf(v[k]) || v[k] < 0.0f) return 0;
    return 1;
}

typedef struct {
    const float* points;
    const float* normals;
    const int32_t* faces;
    const int32_t* nodes;
    const float* bounds;
    const float* materials;
    const float* rgba;
    const float* lum;
    const float* sat;
    int32_t ew, eh;
    const int32_t* light_faces;
    int32_t nlights;
    const float* cam; /* ox oy oz fx fy fz rx ry rz ux uy uz tan_half aspect */
    int32_t w, h, spp, max_bounces, rr_start, flags;
    uint32_t seed;
    float* L;          /* float4 per path */
    int64_t ext, shadow;
    /* light-grid guided environment (lg_mode 1: reference pdf, 2: with the
     * tile pick's pdf; 0: full-map sampling) */
    const float* lg_func;
    const float* lg_sat;
    const float* lg_cam;
    uint64_t* pend;    /* fixed-point accumulation, atomic adds */
    int32_t lg_mode;
} orc_job;

static void lg_add(const orc_job* J, int32_t gidx, float lum)
{ /* accumLightGrid, Light.cl:35-38, in fixed point */
    uint64_t q = orc_lg_fixed(lum);
    if (q) __atomic_fetch_add(&J->pend[gidx], q, __ATOMIC_RELAXED);
}

static int emitter(const float* m)
{
    float s = m[32];
    return luminance3(m[28] * s, m[29] * s, m[30] * s) != 0.0f;
}

typedef struct {
    f4 p1, e1, e2;
    f4 mz;
    float inten[3];
    float inv_area;
} oarea;

static void area_light(const orc_job* J, int tri, oarea* A)
{ /* getAreaLight, Light.cl:43-77 */
    const int32_t* F = J->faces + 10 * (size_t)tri;
    f4 p1 = f4_load(J->points + 4 * (size_t)F[0]);
    f4 p2 = f4_load(J->points + 4 * (size_t)F[1]);
    f4 p3 = f4_load(J->points + 4 * (size_t)F[2]);
    p1.w = p2.w = p3.w = 0.0f;
    f4 n;
    if (F[6] > -1 && J->normals) {
        f4 a = f4_load(J->normals + 4 * (size_t)F[6]), b = f4_load(J->normals + 4 * (size_t)F[7]),
           c = f4_load(J->normals + 4 * (size_t)F[8]);
        f4 s = f4_add(f4_add(a, b), c);
        n = normalize3(f4_make(s.x / 3.0f, s.y / 3.0f, s.z / 3.0f, 0));
    } else {
        n = normalize3(cross3(f4_sub(p2, p1), f4_sub(p3, p1)));
    }
    f4 e1 = f4_sub(p2, p1), e2 = f4_sub(p3, p1);
    A->inv_area = 2.0f / length3(cross3(e1, e2));
    A->mz = n; /* get_frame(n).mZ == n */
    A->p1 = p1;
    A->e1 = e1;
    A->e2 = e2;
    const float* m = J->materials + 36 * (size_t)(F[9] & 0xFFFF);
    float pw = m[33];
    A->inten[0] = m[28] * pw; A->inten[1] = m[29] * pw; A->inten[2] = m[30] * pw;
}

/* Path-id convention of the bounce loop: pid = k * spp + sample, k the
 * pixel's rank when the image is walked in bands of 8 rows, column by
 * column inside a band (the last band holds h % 8 rows). */
void orc_path_pixel(int32_t pid, int32_t w, int32_t h, int32_t spp, int32_t* x, int32_t* y)
{
    int32_t k = pid / spp;
    int32_t full = (h / 8) * 8 * w, band, rows, r;
    if (k < full) {
        band = k / (8 * w);
        rows = 8;
        r = k - band * 8 * w;
    } else {
        band = h / 8;
        rows = h - 8 * band;
        r = k - full;
    }
    *x = r / rows;
    *y = band * 8 + r % rows;
}

static void trace_path(orc_job* J, int32_t pid)
{
    int32_t px, py;
    orc_path_pixel(pid, J->w, J->h, J->spp, &px, &py);
    const float* C = J->cam;
    int32_t st[2];
    seed_for(J->seed, pid, 0, 0, st);
    float j0 = orc_get_random(st), j1 = orc_get_random(st);
    float a = ((((2.0f * ((float)px + j0)) / (float)J->w) - 1.0f) * C[12]) * C[13];
    float b = (1.0f - (2.0f * ((float)py + j1)) / (float)J->h) * C[12];
    f4 dir = normalize3(f4_make((C[3] + a * C[6]) + b * C[9], (C[4] + a * C[7]) + b * C[10],
                                (C[5] + a * C[8]) + b * C[11], 0));
    float o[4] = {C[0], C[1], C[2], 0.001f};
    float d[4] = {dir.x, dir.y, dir.z, INFINITY};
    float T[3] = {1, 1, 1}, last_pdf = 1.0f, L[3] = {0, 0, 0};
    for (int bnc = 0; bnc < J->max_bounces; bnc++) {
        int32_t prim;
        float t;
        J->ext++;
        orc_trace_closest(J->points, J->faces, J->nodes, J->bounds, 1, o, d, &prim, &t, NULL, NULL, 1);
        int last_spec = bnc == 0;
        float add[3] = {0, 0, 0};
        int alive = 0, has_shadow = 0;
        float so[4], sd[4], sc[3];
        int32_t sh_gidx = -1;
        float sh_lum = 0.0f;
        if (prim < 0) {
            if (J->ew > 0) {
                float dd[4] = {d[0], d[1], d[2], 0}, c4[4], pw;
                if (J->lg_mode) { /* getRadianceEnvironmentLight1 at the path's last hit point */
                    float hp[4] = {0, 0, 0, 0}, l;
                    if (bnc > 0) { hp[0] = o[0]; hp[1] = o[1]; hp[2] = o[2]; }
                    int32_t gidx, cell, tile;
                    orc_env_rad_adaptive_one(J->rgba, J->lum, J->sat, J->ew, J->eh, J->lg_cam, dd, hp, c4, &pw, &l,
                                             &gidx, &cell, &tile);
                    if (J->lg_mode == 2) pw = pw * orc_lg_tile_pdf(J->lg_func, J->lg_sat, cell, tile % 16, tile / 16);
                    lg_add(J, gidx, l);
                } else {
                    int32_t tx;
                    orc_env_radiance(J->rgba, J->lum, J->sat, J->ew, J->eh, 1, dd, c4, &pw, &tx);
                }
                if (!(c4[0] == 0.0f && c4[1] == 0.0f && c4[2] == 0.0f)) {
                    float w = last_spec ? 1.0f : mis2(last_pdf, pw);
                    float v[3] = {T[0] * c4[0] * w, T[1] * c4[1] * w, T[2] * c4[2] * w};
                    if (rgb_valid(v)) memcpy(add, v, sizeof(add));
                }
            }
        } else {
            float p[4], n[4], ng[4], uv[2];
            int32_t mat;
            orc_shade_hits(J->points, J->normals, NULL, J->faces, 1, o, d, &prim, &t, p, n, ng, uv, &mat);
            float rec[ORC_BSDF_WORDS];
            int32_t mid;
            float d4[4] = {d[0], d[1], d[2], 0};
            orc_bsdf_setup(1, d4, n, ng, &mat, J->materials, rec, &mid);
            if (mid >= 0) {
                const float* M = J->materials + 36 * (size_t)mat;
                if (emitter(M)) { /* LightHitPass emitter branch, Render.cl:76-104 */
                    oarea A;
                    area_light(J, prim, &A);
                    f4 nd = f4_make(-d[0], -d[1], -d[2], 0);
                    float cos_out = fmaxf(0.0f, dot3(A.mz, nd));
                    if (cos_out != 0.0f && !(A.inten[0] == 0.0f && A.inten[1] == 0.0f && A.inten[2] == 0.0f)) {
                        float w = 1.0f;
                        if (!last_spec) {
                            float pick = 1.0f / (float)J->nlights;
                            float pdfw = (A.inv_area * (t * t)) / fabsf(rec[11]);
                            w = mis2(last_pdf, pdfw * pick);
                        }
                        float v[3] = {T[0] * A.inten[0] * w, T[1] * A.inten[1] * w, T[2] * A.inten[2] * w};
                        if (rgb_valid(v)) memcpy(add, v, sizeof(add));
                    }
                } else {
                    int32_t ls[2];
                    seed_for(J->seed, pid, bnc, 1, ls);
                    float s0 = orc_get_random(ls), s1 = orc_get_random(ls);
                    float s2 = orc_get_random(ls), s3 = orc_get_random(ls);
                    int32_t env_gidx = -1;
                    float env_lum = 0.0f;
                    float pick = 1.0f;
                    int use_env = J->ew > 0, use_mesh = 0;
                    if (use_env && J->nlights > 0) {
                        float r = orc_get_random(ls);
                        pick *= 0.5f;
                        use_mesh = r < 0.5f;
                        use_env = !use_mesh;
                    } else if (!use_env && J->nlights > 0) {
                        use_mesh = 1;
                    }
                    float ldir[4] = {0, 0, 0, 0}, dist = 0.0f, lpdf = 0.0f, rad[4] = {0, 0, 0, 0};
                    if (use_env && J->lg_mode) { /* illuminateEnvironmentLight1 */
                        float q4[4] = {s0, s1, s2, s3}, tpdf, p4[4] = {p[0], p[1], p[2], 0};
                        orc_env_illum_adaptive_one(J->rgba, J->lum, J->sat, J->ew, J->eh, J->lg_func, J->lg_sat,
                                                   J->lg_cam, p4, q4, ldir, rad, &lpdf, &env_lum, &env_gidx, NULL,
                                                   &tpdf);
                        if (J->lg_mode == 2) lpdf = (tpdf > 0.0f && tpdf < INFINITY) ? lpdf * tpdf : 0.0f;
                        dist = 3.402823e+38f;
                    } else if (use_env) {
                        float u2[2] = {s0, s1};
                        int32_t tx;
                        orc_env_illuminate(J->rgba, J->lum, J->sat, J->ew, J->eh, 1, u2, ldir, rad, &lpdf, &tx);
                        dist = 3.402823e+38f;
                    } else if (use_mesh) {
                        int li = (int)(orc_get_random(ls) * (float)J->nlights);
                        int tri = (J->flags & 1) ? J->light_faces[li] : J->light_faces[0];
                        pick *= 1.0f / (float)J->nlights;
                        oarea A;
                        area_light(J, tri, &A);
                        float su = sqrtf(s0);
                        float bu = 1.0f - su, bv = s1 * su;
                        f4 lp = f4_add(f4_add(A.p1, f4_scale(A.e1, bu)), f4_scale(A.e2, bv));
                        f4 dv = f4_make(lp.x - p[0], lp.y - p[1], lp.z - p[2], 0);
                        float dist2 = dot3(dv, dv);
                        dist = sqrtf(dist2);
                        ldir[0] = dv.x / dist; ldir[1] = dv.y / dist; ldir[2] = dv.z / dist;
                        float cn = dot3(A.mz, f4_make(-ldir[0], -ldir[1], -ldir[2], 0));
                        if (!(cn < 1e-6f)) {
                            lpdf = (A.inv_area * dist2) / cn;
                            rad[0] = A.inten[0]; rad[1] = A.inten[1]; rad[2] = A.inten[2];
                        }
                    }
                    if (!(rad[0] == 0.0f && rad[1] == 0.0f && rad[2] == 0.0f)) {
                        float f[3], bpdf, cosg;
                        float l3[3] = {ldir[0], ldir[1], ldir[2]};
                        orc_bsdf_eval(1, rec, &mid, J->materials, l3, f, &bpdf, &cosg);
                        if (!(f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f)) {
                            float w = mis2(lpdf, bpdf);
                            float s = (w * cosg) / (pick * lpdf);
                            float cv[3] = {s * (rad[0] * f[0]) * T[0], s * (rad[1] * f[1]) * T[1],
                                           s * (rad[2] * f[2]) * T[2]};
                            if (rgb_valid(cv)) {
                                so[0] = p[0]; so[1] = p[1]; so[2] = p[2]; so[3] = 0.0001f;
                                sd[0] = ldir[0]; sd[1] = ldir[1]; sd[2] = ldir[2]; sd[3] = dist - 2 * 1e-3f;
                                memcpy(sc, cv, sizeof(sc));
                                has_shadow = 1;
                                sh_gidx = env_gidx;
                                sh_lum = env_lum;
                            }
                        }
                    }
                    int32_t bs[2];
                    seed_for(J->seed, pid, bnc, 2, bs);
                    float u[3];
                    u[0] = orc_get_random(bs); u[1] = orc_get_random(bs); u[2] = orc_get_random(bs);
                    float f[3], wd[3], pdf, cosg;
                    orc_bsdf_sample(1, rec, &mid, J->materials, u, f, wd, &pdf, &cosg);
                    float k = cosg / pdf;
                    float Tn[3] = {T[0] * (f[0] * k), T[1] * (f[1] * k), T[2] * (f[2] * k)};
                    alive = pdf > 0.0f && !(f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f) && rgb_valid(Tn) &&
                            !(Tn[0] == 0.0f && Tn[1] == 0.0f && Tn[2] == 0.0f);
                    if (alive && J->rr_start >= 0 && bnc + 1 >= J->rr_start) {
                        float q = fminf(1.0f, fmaxf(Tn[0], fmaxf(Tn[1], Tn[2])));
                        float r = orc_get_random(bs);
                        if (r >= q) alive = 0;
                        else { Tn[0] = Tn[0] / q; Tn[1] = Tn[1] / q; Tn[2] = Tn[2] / q; }
                    }
                    if (alive) {
                        memcpy(T, Tn, sizeof(T));
                        last_pdf = pdf;
                        o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = 0.001f;
                        d[0] = wd[0]; d[1] = wd[1]; d[2] = wd[2]; d[3] = INFINITY;
                    }
                }
            }
        }
        if (add[0] != 0.0f || add[1] != 0.0f || add[2] != 0.0f) {
            L[0] = L[0] + add[0]; L[1] = L[1] + add[1]; L[2] = L[2] + add[2];
        }
        if (has_shadow) {
            int32_t occ;
            J->shadow++;
            orc_trace_any(J->points, J->faces, J->nodes, J->bounds, 1, so, sd, &occ, NULL, NULL, 1);
            if (!occ) {
                L[0] = L[0] + sc[0]; L[1] = L[1] + sc[1]; L[2] = L[2] + sc[2];
                if (J->lg_mode && sh_gidx >= 0) lg_add(J, sh_gidx, sh_lum); /* Render.cl:458-459 */
            }
        }
        if (!alive) break;
    }
    float* out = J->L + 4 * (size_t)pid;
    out[0] = L[0]; out[1] = L[1]; out[2] = L[2]; out[3] = 0.0f;
}

typedef struct {
    orc_job* J;
    const int32_t* pids;
    int64_t begin, end;
    int64_t ext, shadow;
} orc_rw;

static void* render_worker(void* arg)
{
    orc_rw* W = (orc_rw*)arg;
    orc_job J = *W->J;
    J.ext = 0;
    J.shadow = 0;
    for (int64_t i = W->begin; i < W->end; i++) trace_path(&J, W->pids[i]);
    W->ext = J.ext;
    W->shadow = J.shadow;
    return NULL;
}

/* Pinhole basis as the product's host code computes it (double, then float). */
int orc_camera_basis(const float* pos, const float* look, const float* up, float vfov_deg,
                     int32_t w, int32_t h, float* cam)
{
    double f[3] = {(double)look[0] - pos[0], (double)look[1] - pos[1], (double)look[2] - pos[2]};
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int k = 0; k < 3; k++) f[k] /= fl;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int k = 0; k < 3; k++) r[k] /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    cam[0] = pos[0]; cam[1] = pos[1]; cam[2] = pos[2];
    for (int k = 0; k < 3; k++) { cam[3 + k] = (float)f[k]; cam[6 + k] = (float)r[k]; cam[9 + k] = (float)u[k]; }
    cam[12] = (float)tan(0.5 * vfov_deg * 3.14159265358979323846 / 180.0);
    cam[13] = (float)w / (float)h;
    return ORC_OK;
}

/* Traces the listed path ids; L receives float4 per path id (caller-sized to
 * the largest id + 1).  cam: ox oy oz fx fy fz rx ry rz ux uy uz tan_half
 * aspect, as the product computes them.  rays[2] = extension, shadow. */
int orc_render_paths(const float* points, const float* normals, const int32_t* faces,
                     const int32_t* nodes, const float* bounds, const float* materials,
                     const float* rgba, const float* lum, const float* sat, int32_t ew, int32_t eh,
                     const int32_t* light_faces, int32_t nlights, const float* cam, int32_t w,
                     int32_t h, int32_t spp, int32_t max_bounces, int32_t rr_start, uint32_t seed,
                     int32_t flags, int64_t npaths, const int32_t* pids, float* L, int64_t* rays,
                     int32_t nthreads)
{
    return orc_render_paths_lg(points, normals, faces, nodes, bounds, materials, rgba, lum, sat, ew, eh,
                               light_faces, nlights, cam, w, h, spp, max_bounces, rr_start, seed, flags, npaths,
                               pids, L, rays, nthreads, NULL, NULL, NULL, NULL, 0);
}

/* orc_render_paths with the light-grid guided environment: lg_func / lg_sat
 * the grid's arrays (1600 x 1600), lg_cam its camera position, pend the
 * fixed-point accumulation (uint64[1600 * 1600], added to atomically),
 * lg_mode 1 (the reference's in-tile pdf) or 2 (tile pick's pdf applied). */
int orc_render_paths_lg(const float* points, const float* normals, const int32_t* faces,
                        const int32_t* nodes, const float* bounds, const float* materials,
                        const float* rgba, const float* lum, const float* sat, int32_t ew, int32_t eh,
                        const int32_t* light_faces, int32_t nlights, const float* cam, int32_t w,
                        int32_t h, int32_t spp, int32_t max_bounces, int32_t rr_start, uint32_t seed,
                        int32_t flags, int64_t npaths, const int32_t* pids, float* L, int64_t* rays,
                        int32_t nthreads, const float* lg_func, const float* lg_sat, const float* lg_cam,
                        uint64_t* pend, int32_t lg_mode)
{
    orc_job J;
    memset(&J, 0, sizeof(J));
    if (lg_mode && !(lg_func && lg_sat && lg_cam && pend && rgba)) return ORC_ERR_ARG;
    J.lg_func = lg_func; J.lg_sat = lg_sat; J.lg_cam = lg_cam; J.pend = pend; J.lg_mode = lg_mode;
    J.points = points; J.normals = normals; J.faces = faces; J.nodes = nodes; J.bounds = bounds;
    J.materials = materials; J.rgba = rgba; J.lum = lum; J.sat = sat; J.ew = rgba ? ew : 0; J.eh = eh;
    J.light_faces = light_faces; J.nlights = light_faces ? nlights : 0; J.cam = cam;
    J.w = w; J.h = h; J.spp = spp; J.max_bounces = max_bounces; J.rr_start = rr_start;
    J.flags = flags; J.seed = seed; J.L = L;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    orc_rw W[256];
    pthread_t th[256];
    int64_t chunk = (npaths + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; k++) {
        W[k].J = &J;
        W[k].pids = pids;
        W[k].begin = k * chunk < npaths ? k * chunk : npaths;
        W[k].end = (k + 1) * chunk < npaths ? (k + 1) * chunk : npaths;
        W[k].ext = W[k].shadow = 0;
    }
    if (nthreads == 1) render_worker(&W[0]);
    else {
        for (int k = 0; k < nthreads; k++) pthread_create(&th[k], NULL, render_worker, &W[k]);
        for (int k = 0; k < nthreads; k++) pthread_join(th[k], NULL);
    }
    int64_t e = 0, s = 0;
    for (int k = 0; k < nthreads; k++) { e += W[k].ext; s += W[k].shadow; }
    if (rays) { rays[0] = e; rays[1] = s; }
    return ORC_OK;
}
