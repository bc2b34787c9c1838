/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain-C CPU restatement of the reference's wavefront path-tracing core
 * (mambastudio/MambaTracer, the OpenCL kernels under src/cl/kernel and the
 * Java host drivers).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline.  The product
 * (mambatracer_amd/libmambart.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Compiled with
 * -O3 -ffp-contract=off (no FMA contraction), IEEE float, so the HIP kernels,
 * which follow the same evaluation order, can be compared bit for bit on the
 * integer outputs (primitive ids, compaction orders, light-grid cells).
 *
 * Parity status of each part is listed in the header comment of the file that
 * implements it.  Parts whose reference arithmetic lives in an un-vendored jar
 * (BVH partition, host env SAT, light-grid seeding) are marked
 * "parity unpinned" there.
 *
 * Layouts (the reference's flat layout, src/cl/kernel/Util.cl:95-121,
 * NormalBVH.cl:1-11):
 *   points  float[4*V]    (x,y,z,w=0)
 *   faces   int32[10*N]   {v0,v1,v2, uv0,uv1,uv2, n0,n1,n2, mat}
 *   nodes   int32[7*(2N-1)] {bound, sibling, left, right, parent, isLeaf, child}
 *   bounds  float[8*(2N-1)] {min.xyzw, max.xyzw}
 * Rays are passed as two float4 streams: (o.xyz, tMin) and (d.xyz, tMax).
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_DEPTH (-2)

/* ---- A: BVH build and traversal ---- */
int orc_bvh_build(int32_t ntris, const float* points, const int32_t* faces,
                  int32_t* nodes, float* bounds, int32_t* max_depth);

int orc_trace_closest(const float* points, const int32_t* faces,
                      const int32_t* nodes, const float* bounds,
                      int64_t nrays, const float* o_tmin, const float* d_tmax,
                      int32_t* prim, float* t, int32_t* ray_stats,
                      int64_t* total_stats, int32_t nthreads);

int orc_trace_any(const float* points, const int32_t* faces,
                  const int32_t* nodes, const float* bounds,
                  int64_t nrays, const float* o_tmin, const float* d_tmax,
                  int32_t* occluded, int32_t* ray_stats,
                  int64_t* total_stats, int32_t nthreads);

int orc_trace_brute(const float* points, const int32_t* faces, int32_t ntris,
                    int64_t nrays, const float* o_tmin, const float* d_tmax,
                    int32_t* prim, float* t);

int orc_shade_hits(const float* points, const float* normals, const float* uvs,
                   const int32_t* faces, int64_t nrays, const float* o_tmin,
                   const float* d_tmax, const int32_t* prim, const float* t,
                   float* p, float* n, float* ng, float* uv, int32_t* mat);

int orc_shadow_rays(int64_t n, const float* o_tmin, const float* d_tmax,
                    const int32_t* prim, const float* t, const float* light,
                    float* s_o_tmin, float* s_d_tmax, int32_t* valid);

/* ---- D: scan and compaction ---- */
int orc_scan_excl_i32(const int32_t* in, int64_t n, int32_t* out, int32_t* total);
int orc_compact_stable(const int32_t* pred, int64_t n, int32_t* out_idx, int32_t* count);
int32_t orc_scan_excl_f32(const float* in, int64_t n, float* out, float* total);

/* ---- image pipeline (CImage.processImage) ---- */
int32_t orc_image_process(const float* accum, float frame_count, int32_t n, float exposure, float gamma,
                          int32_t update_totals, float* frame, int32_t* argb_io, float* total_log,
                          int32_t* total_count);
int32_t orc_image_init_argb(void);
int32_t orc_argb(float r, float g, float b, float a);

/* ---- C: light grid, SATs, environment light ---- */
/* shared transcendentals (mt_xmath.h): kind 0 sin, 1 cos, 2 acos, 3 atan2(a, b), 4 a^5 */
int orc_xm_eval(int32_t kind, int64_t n, const float* a, const float* b, float* out);
int orc_sat_tiles(const float* func, float* sat, int32_t nu, int32_t nv,
                  int32_t rx, int32_t ry);
int orc_lg_set_tile_func(float* func, const float* tile);
int orc_lg_sample(const float* func, const float* sat, int32_t n,
                  const int32_t* cell, const float* u, int32_t* tile_xy,
                  float* uv, float* pdf);
int orc_lg_cell_from_dirs(int32_t n, const float* dirs, int32_t* cell, int32_t* tile_xy);

int orc_env_lum(const float* rgba, int32_t w, int32_t h, float* lum);
int orc_env_sat(const float* lum, int32_t w, int32_t h, float* sat);
int orc_sat_sample_region(const float* func, const float* sat, int32_t nu, int32_t nv,
                          int32_t minx, int32_t miny, int32_t sx, int32_t sy,
                          int32_t n, const float* u, float* uv, int32_t* offset, float* pdf);
int orc_env_illuminate(const float* rgba, const float* lum, const float* sat,
                       int32_t w, int32_t h, int32_t n, const float* u,
                       float* dir, float* contrib, float* pdfw, int32_t* texel);
int orc_env_radiance(const float* rgba, const float* lum, const float* sat,
                     int32_t w, int32_t h, int32_t n, const float* dir,
                     float* contrib, float* pdfw, int32_t* texel);
int orc_env_illuminate_adaptive(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                                const float* lg_func, const float* lg_sat_arr, const float* cam, int32_t n,
                                const float* hit, const float* u, float* dir, float* contrib, float* pdfw,
                                float* lum_out, int32_t* gidx, int32_t* texel_xy);
int orc_env_radiance_adaptive(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                              float* accum, const float* cam, int32_t n, const float* dir, const float* hit,
                              float* contrib, float* pdfa, float* lum_out, int32_t* gidx);

/* one-sample forms used by the bounce loop, the tile-pick pdf and the
 * fixed-point grid accumulation of the guided environment mode */
void orc_env_illum_adaptive_one(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                                const float* lg_func, const float* lg_sat_arr, const float* cam, const float* hit,
                                const float* q, float* dir, float* contrib, float* pdfw, float* lum_out,
                                int32_t* gidx, int32_t* texel_xy, float* tile_pdf);
void orc_env_rad_adaptive_one(const float* rgba, const float* lum, const float* sat, int32_t w, int32_t h,
                              const float* cam, const float* dir, const float* hit, float* contrib, float* pdfa,
                              float* lum_out, int32_t* gidx, int32_t* cell, int32_t* tile);
float orc_lg_tile_pdf(const float* lg_func, const float* lg_sat_arr, int32_t s, int32_t tx, int32_t ty);
uint64_t orc_lg_fixed(float lum);
int orc_lg_fold(uint64_t* pend, float* accum);

/* ---- B: BSDF ---- */
/* materials: 36 x 4-byte words per material (4 int flags + 8 float4). */
#define ORC_BSDF_WORDS 20
int orc_bsdf_setup(int32_t n, const float* ray_d, const float* isect_n,
                   const float* isect_ng, const int32_t* mat,
                   const float* materials, float* bsdf, int32_t* mat_id);
int orc_bsdf_sample(int32_t n, const float* bsdf, const int32_t* mat_id,
                    const float* materials, const float* u, float* f,
                    float* dir, float* pdf, float* cos_out);
int orc_bsdf_eval(int32_t n, const float* bsdf, const int32_t* mat_id,
                  const float* materials, const float* dir, float* f,
                  float* pdf, float* cos_out);
float orc_ggx_D(const float* wh, float ax, float ay);
float orc_ggx_lambda(const float* w, float ax, float ay);
void orc_ggx_sample_vndf(const float* v, float ax, float ay, float u1, float u2, float* out);
float orc_ggx_pdf_vndf(const float* v, const float* h, float ax, float ay);

/* ---- bounce loop (CDeviceGI.loop + Render.cl kernels), one path at a time ---- */
/* pixel of path `pid`: pid = rank * spp + sample, pixels ranked in bands of 8
 * rows, column by column inside a band (the last band holds h % 8 rows) */
void orc_path_pixel(int32_t pid, int32_t w, int32_t h, int32_t spp, int32_t* x, int32_t* y);
int orc_camera_basis(const float* pos, const float* look, const float* up, float vfov_deg,
                     int32_t w, int32_t h, float* cam);
int orc_render_paths(const float* points, const float* normals, const int32_t* faces,
                     const int32_t* nodes, const float* bounds, const float* materials,
                     const float* rgba, const float* lum, const float* sat, int32_t ew, int32_t eh,
                     const int32_t* light_faces, int32_t nlights, const float* cam, int32_t w,
                     int32_t h, int32_t spp, int32_t max_bounces, int32_t rr_start, uint32_t seed,
                     int32_t flags, int64_t npaths, const int32_t* pids, float* L, int64_t* rays,
                     int32_t nthreads);
int orc_render_paths_lg(const float* points, const float* normals, const int32_t* faces,
                        const int32_t* nodes, const float* bounds, const float* materials,
                        const float* rgba, const float* lum, const float* sat, int32_t ew, int32_t eh,
                        const int32_t* light_faces, int32_t nlights, const float* cam, int32_t w,
                        int32_t h, int32_t spp, int32_t max_bounces, int32_t rr_start, uint32_t seed,
                        int32_t flags, int64_t npaths, const int32_t* pids, float* L, int64_t* rays,
                        int32_t nthreads, const float* lg_func, const float* lg_sat, const float* lg_cam,
                        uint64_t* pend, int32_t lg_mode);

/* ---- RNG (Sampling.cl:15-59) ---- */
uint32_t orc_wang_hash(int32_t seed);
float orc_get_random(int32_t* state2);

#ifdef __cplusplus
}
#endif
#endif
