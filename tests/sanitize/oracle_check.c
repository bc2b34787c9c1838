/* The oracle (oracle/*.c) driven over small scenes through every stage:
 * BVH build, closest / any / brute-force hits, hit shading, scans and
 * compaction, tile SATs, light-grid and environment sampling (full-map and
 * guided), BSDF setup / sample / eval, the bounce loop in every lighting
 * mode with mesh lights, the image pipeline and the shared transcendentals.
 * Built with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitize.py; exit status 0 when every call returned OK. */
#include "../../oracle/orc.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rs = 88172645463325252ull;
static float frand(void)
{
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return (float)((rs >> 40) & 0xffffff) / 16777216.0f;
}

#define CHECK(x)                                                         \
    do {                                                                 \
        if ((x) != ORC_OK) {                                             \
            fprintf(stderr, "%s failed at line %d\n", #x, __LINE__);     \
            return 1;                                                    \
        }                                                                \
    } while (0)

int main(void)
{
    const int ntri = 3000, nv = 3 * ntri;
    float* pts = calloc((size_t)nv * 4, sizeof(float));
    int32_t* faces = malloc((size_t)ntri * 10 * sizeof(int32_t));
    for (int i = 0; i < ntri; i++) {
        float c[3] = {2 * frand() - 1, 2 * frand() - 1, 2 * frand() - 1};
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) pts[4 * (3 * i + k) + a] = c[a] + 0.1f * (frand() - 0.5f);
        for (int k = 0; k < 10; k++) faces[10 * i + k] = -1;
        faces[10 * i] = 3 * i; faces[10 * i + 1] = 3 * i + 1; faces[10 * i + 2] = 3 * i + 2;
        faces[10 * i + 9] = i % 4;
    }
    int32_t* nodes = malloc((size_t)(2 * ntri - 1) * 7 * sizeof(int32_t));
    float* bounds = malloc((size_t)(2 * ntri - 1) * 8 * sizeof(float));
    int32_t depth = 0;
    CHECK(orc_bvh_build(ntri, pts, faces, nodes, bounds, &depth));

    const int nr = 4000;
    float* o = malloc((size_t)nr * 4 * sizeof(float));
    float* d = malloc((size_t)nr * 4 * sizeof(float));
    for (int i = 0; i < nr; i++) {
        float v[3] = {frand() - 0.5f, frand() - 0.5f, frand() - 0.5f};
        float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1e-6f;
        for (int a = 0; a < 3; a++) { o[4 * i + a] = 3 * frand() - 1.5f; d[4 * i + a] = v[a] / l; }
        o[4 * i + 3] = 0.001f;
        d[4 * i + 3] = i % 5 ? INFINITY : 1.0f;
    }
    int32_t* prim = malloc(nr * sizeof(int32_t));
    int32_t* prim2 = malloc(nr * sizeof(int32_t));
    int32_t* occ = malloc(nr * sizeof(int32_t));
    int32_t* st = malloc((size_t)nr * 3 * sizeof(int32_t));
    float* t = malloc(nr * sizeof(float));
    float* t2 = malloc(nr * sizeof(float));
    int64_t tot[3];
    CHECK(orc_trace_closest(pts, faces, nodes, bounds, nr, o, d, prim, t, st, tot, 3));
    CHECK(orc_trace_any(pts, faces, nodes, bounds, nr, o, d, occ, st, tot, 2));
    CHECK(orc_trace_brute(pts, faces, ntri, 200, o, d, prim2, t2));
    float* P = malloc((size_t)nr * 4 * sizeof(float));
    float* N = malloc((size_t)nr * 4 * sizeof(float));
    float* NG = malloc((size_t)nr * 4 * sizeof(float));
    float* UV = malloc((size_t)nr * 2 * sizeof(float));
    int32_t* mat = malloc(nr * sizeof(int32_t));
    CHECK(orc_shade_hits(pts, NULL, NULL, faces, nr, o, d, prim, t, P, N, NG, UV, mat));

    int32_t* pred = malloc(nr * sizeof(int32_t));
    int32_t* scan = malloc(nr * sizeof(int32_t));
    int32_t total, count;
    for (int i = 0; i < nr; i++) pred[i] = prim[i] >= 0;
    CHECK(orc_scan_excl_i32(pred, nr, scan, &total));
    CHECK(orc_compact_stable(pred, nr, scan, &count));
    float* fs = malloc(nr * sizeof(float));
    float fsum;
    for (int i = 0; i < nr; i++) UV[i] = frand();
    CHECK(orc_scan_excl_f32(UV, nr, fs, &fsum));

    /* light grid */
    const size_t LG = 1600u * 1600u;
    float* func = calloc(LG, sizeof(float));
    float* sat = calloc(LG, sizeof(float));
    float tile[512];
    for (int i = 0; i < 512; i++) tile[i] = i % 7 ? frand() : 0.0f;
    CHECK(orc_lg_set_tile_func(func, tile));
    CHECK(orc_sat_tiles(func, sat, 1600, 1600, 16, 32));
    int32_t cell[64], txy[128];
    float u2[128], uvo[128], pdf[64];
    for (int i = 0; i < 64; i++) { cell[i] = (i * 97) % 5000; u2[2 * i] = frand(); u2[2 * i + 1] = frand(); }
    CHECK(orc_lg_sample(func, sat, 64, cell, u2, txy, uvo, pdf));
    CHECK(orc_lg_cell_from_dirs(64, d, cell, txy));

    /* environment */
    const int ew = 64, eh = 32;
    float* rgba = malloc((size_t)ew * eh * 4 * sizeof(float));
    for (int i = 0; i < ew * eh; i++) for (int a = 0; a < 4; a++) rgba[4 * i + a] = 0.2f + frand();
    float* lum = malloc((size_t)ew * eh * sizeof(float));
    float* esat = malloc((size_t)ew * eh * sizeof(float));
    CHECK(orc_env_lum(rgba, ew, eh, lum));
    CHECK(orc_env_sat(lum, ew, eh, esat));
    float dir4[256], con4[256], pw[64], lo[64];
    int32_t tx[64], gi[64], tex2[128];
    CHECK(orc_env_illuminate(rgba, lum, esat, ew, eh, 64, u2, dir4, con4, pw, tx));
    CHECK(orc_env_radiance(rgba, lum, esat, ew, eh, 64, d, con4, pw, tx));
    float u4[256];
    for (int i = 0; i < 256; i++) u4[i] = frand();
    const float cam[3] = {0.3f, 0.2f, -3.0f};
    CHECK(orc_env_illuminate_adaptive(rgba, lum, esat, ew, eh, func, sat, cam, 64, o, u4, dir4, con4, pw, lo, gi,
                                      tex2));
    float* accum = calloc(LG, sizeof(float));
    CHECK(orc_env_radiance_adaptive(rgba, lum, esat, ew, eh, accum, cam, 64, d, o, con4, pw, lo, gi));

    /* BSDF */
    float mats[4 * 36];
    memset(mats, 0, sizeof(mats));
    for (int m = 0; m < 4; m++) {
        float* M = mats + 36 * m;
        M[4] = M[5] = M[6] = 0.5f;
        M[8] = 1.0f;
        M[12] = M[13] = M[14] = 0.04f;
        M[16] = 1.0f;
        M[17] = 0.1f + 0.2f * m;
        M[18] = 0.3f;
        M[28] = M[29] = M[30] = 1.0f;
    }
    mats[36 * 3 + 32] = 1.0f; /* material 3 emits */
    mats[36 * 3 + 33] = 4.0f;
    float* rec = malloc((size_t)nr * ORC_BSDF_WORDS * sizeof(float));
    int32_t* mid = malloc(nr * sizeof(int32_t));
    CHECK(orc_bsdf_setup(nr, d, N, NG, mat, mats, rec, mid));
    float* u3 = malloc((size_t)nr * 3 * sizeof(float));
    for (int i = 0; i < 3 * nr; i++) u3[i] = frand();
    float* f3 = malloc((size_t)nr * 3 * sizeof(float));
    float* d3 = malloc((size_t)nr * 3 * sizeof(float));
    float* cs = malloc(nr * sizeof(float));
    CHECK(orc_bsdf_sample(nr, rec, mid, mats, u3, f3, d3, t2, cs));
    CHECK(orc_bsdf_eval(nr, rec, mid, mats, d3, f3, t2, cs));

    /* bounce loop: full map, guided (both pdfs), mesh lights, no light */
    float cb[14];
    const float pos[3] = {0, 0, -3}, look[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    CHECK(orc_camera_basis(pos, look, up, 50.0f, 24, 16, cb));
    const int np = 24 * 16 * 2;
    int32_t* pids = malloc(np * sizeof(int32_t));
    for (int i = 0; i < np; i++) pids[i] = i;
    float* L = malloc((size_t)np * 4 * sizeof(float));
    int64_t rays[2];
    int32_t lights[ntri];
    int32_t nl = 0;
    for (int i = 0; i < ntri; i++) if (faces[10 * i + 9] == 3) lights[nl++] = i;
    uint64_t* pend = calloc(LG, sizeof(uint64_t));
    for (int mode = 0; mode < 3; mode++)
        CHECK(orc_render_paths_lg(pts, NULL, faces, nodes, bounds, mats, rgba, lum, esat, ew, eh, lights, nl, cb,
                                  24, 16, 2, 6, 3, 11, 1, np, pids, L, rays, 3, func, sat, cam, pend, mode));
    CHECK(orc_lg_fold(pend, accum));
    CHECK(orc_render_paths(pts, NULL, faces, nodes, bounds, mats, NULL, NULL, NULL, 0, 0, NULL, 0, cb, 24, 16, 2,
                           6, -1, 12, 0, np, pids, L, rays, 2));

    /* image pipeline and shared transcendentals */
    float* frame = malloc((size_t)np * 4 * sizeof(float));
    int32_t* argb = calloc(np, sizeof(int32_t));
    float tlog = 0.0f;
    int32_t tcnt = 0;
    CHECK(orc_image_process(L, 2.0f, np, 1.0f, 2.2f, 1, frame, argb, &tlog, &tcnt));
    CHECK(orc_xm_eval(3, nr, o, d, t));
    CHECK(orc_xm_eval(2, nr, d, NULL, t));

    free(pts); free(faces); free(nodes); free(bounds); free(o); free(d); free(prim); free(prim2); free(occ);
    free(st); free(t); free(t2); free(P); free(N); free(NG); free(UV); free(mat); free(pred); free(scan);
    free(fs); free(func); free(sat); free(rgba); free(lum); free(esat); free(accum); free(rec); free(mid);
    free(u3); free(f3); free(d3); free(cs); free(pids); free(L); free(pend); free(frame); free(argb);
    printf("oracle check done\n");
    return 0;
}
