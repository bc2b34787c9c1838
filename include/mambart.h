/*
 * mambart — MI355X (gfx950) wavefront path-tracing core, C ABI.
 *
 * Drop-in for the OpenCL kernel launches of mambastudio/MambaTracer's
 * wavefront path tracer.  The Java host today launches each kernel through
 * the un-vendored wrapper.core/JOCL layer
 * (OpenCLConfiguration.createKernel / execute1DKernel, e.g.
 * src/cl/device/CDeviceGI.java:230-276, src/cl/algorithms/CCompact.java:79-89,
 * src/cl/algorithms/PrefixSumInteger.java:80-88,
 * src/cl/algorithms/CAdaptiveEnvironment.java:69-145); each entry point below
 * replaces one such launch sequence and cites it.
 *
 * Conventions (kept from the reference's OpenCL kernels in src/cl/kernel):
 *  - The caller owns every buffer; kernels never allocate caller-visible
 *    memory.  Buffers are device pointers (mt_buf_alloc, or any HIP
 *    allocation on the context's device).
 *  - Counts that flow between launches are passed BY DEVICE POINTER
 *    (`const int32_t* d_count`), as the reference passes `global int* count`
 *    (NormalBVH.cl:97,120), so no host synchronisation is needed between
 *    stages.  `capacity` is the number of slots the buffers hold (the
 *    reference's global work size); *d_count is clamped to it.
 *  - All calls are asynchronous on the context's stream; mt_ctx_sync waits.
 *  - Every function returns 0 (MT_OK) or a negative MT_ERR_* code;
 *    mt_last_error copies a thread-local message.  A Java/Panama binding
 *    turns a non-zero status into an exception, as JOCL's
 *    CL.setExceptionsEnabled(true) does (src/cl/ui/fx/main/TracerAPI.java:79).
 *  - Scene data stays in the reference's flat layout
 *    (src/cl/kernel/Util.cl:95-121, NormalBVH.cl:1-11):
 *      points  float4[V]            (x, y, z, w = 0)
 *      faces   int32[10*N]          {v0,v1,v2, uv0,uv1,uv2, n0,n1,n2, mat}
 *      nodes   int32[7*(2N-1)]      {bound, sibling, left, right, parent, isLeaf, child}
 *      bounds  float[8*(2N-1)]      {min.xyzw, max.xyzw}
 *      materials 36 x 4-byte words  (4 int flags + 8 float4, Bsdf.cl:24-42)
 *  - Rays are two float4 streams (SoA of 16-byte records):
 *      o_tmin[i] = (o.x, o.y, o.z, tMin),  d_tmax[i] = (d.x, d.y, d.z, tMax)
 *    replacing the 80-byte AoS Ray (Util.cl:129-138); inv_d and sign are
 *    recomputed in-kernel exactly as initGlobalRay does (Util.cl:577-600).
 *
 * Thread safety: calls on one context are issued in order from one thread
 * (the reference's one render thread per device, CDeviceGI.java:55);
 * distinct contexts may be used from distinct threads.
 */
#ifndef MAMBART_H
#define MAMBART_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MT_OK 0
#define MT_ERR_ARG (-1)      /* invalid argument / shape */
#define MT_ERR_HIP (-2)      /* HIP runtime error */
#define MT_ERR_ALLOC (-3)    /* device or host allocation failed */
#define MT_ERR_DEPTH (-4)    /* BVH deeper than the traversal supports */
#define MT_ERR_CAPACITY (-5) /* request exceeds a documented maximum */
#define MT_ERR_STATE (-6)    /* object used in the wrong state */

typedef struct mt_context mt_context;
typedef struct mt_scene mt_scene;
typedef struct mt_lightgrid mt_lightgrid;
typedef struct mt_envmap mt_envmap;

/* ------------------------------------------------------------------ */
/* Errors, context, buffers (replace wrapper.core.OpenCLConfiguration / */
/* CMemory: createBufferB/I/F, transferToDevice, finish —               */
/* usage e.g. src/cl/device/CDeviceGI.java:150-179, 428)               */
/* ------------------------------------------------------------------ */
int32_t mt_last_error(char* buf, int32_t len);
const char* mt_version(void);

/* flags: MT_CTX_OWN_STREAM creates a private non-blocking stream and ignores
 * `stream`; otherwise work is issued on `stream` (a hipStream_t; NULL is the
 * device's null stream). */
#define MT_CTX_OWN_STREAM 1
int32_t mt_ctx_create(int32_t device, void* stream, int32_t flags, mt_context** out);
/* Objects made on a context (scenes, light grids, environments, batches)
 * keep a pointer to it: destroy them first. */
int32_t mt_ctx_destroy(mt_context* ctx);
int32_t mt_ctx_sync(mt_context* ctx); /* OpenCLConfiguration.finish() */

int32_t mt_buf_alloc(mt_context* ctx, uint64_t bytes, void** dptr);
int32_t mt_buf_free(mt_context* ctx, void* dptr);
int32_t mt_buf_upload(mt_context* ctx, void* dst, const void* src, uint64_t bytes);
int32_t mt_buf_download(mt_context* ctx, void* dst, const void* src, uint64_t bytes);
int32_t mt_buf_fill_i32(mt_context* ctx, int32_t* dst, int32_t value, uint64_t count);

/* ------------------------------------------------------------------ */
/* Host-side scene preparation                                        */
/* ------------------------------------------------------------------ */

/* CNormalBVH.build (src/cl/scene/CNormalBVH.java:45-148): longest-axis
 * centroid-midpoint split, one primitive per leaf, 2N-1 nodes allocated
 * depth-first with sibling pairs adjacent.  Host arrays; nthreads <= 0
 * uses all cores.  max_depth receives the deepest leaf (root = 0). */
int32_t mt_bvh_build(int32_t ntris, const float* points, const int32_t* faces,
                     int32_t* nodes, float* bounds, int32_t* max_depth,
                     int32_t nthreads);

/* Seeded synthetic inputs (no counterpart in the reference; they make the
 * benchmark workloads).  A "soup" is random triangles whose centroids are
 * uniform in [c_lo, c_hi]^3 and whose two edges have uniformly random
 * directions and lengths uniform in [e_lo, e_hi]; vertex 0 is
 * c - (e1 + e2)/3.  Outputs are host arrays sized by the caller:
 * points float4[3*n], faces int32[10*n] (uv/normal indices -1, material
 * `mat`), with vertex and face numbering starting at `first_vertex`. */
int32_t mt_gen_soup(uint64_t seed, int32_t n, float c_lo, float c_hi,
                    float e_lo, float e_hi, int32_t first_vertex, int32_t mat,
                    float* points, int32_t* faces);
/* The 12-triangle room box [-1,1]^3 with inward-facing triangles:
 * points float4[8], faces int32[10*12]. */
int32_t mt_gen_room(int32_t mat, float* points, int32_t* faces);
/* Pinhole primary rays at pixel centres: origin pos, looking at `look`,
 * up vector `up`, vertical field of view vfov_deg, w x h, tMin 0.001,
 * tMax +inf (initGlobalRay, Util.cl:577-587). */
int32_t mt_gen_camera_rays(const float* pos, const float* look, const float* up,
                           float vfov_deg, int32_t w, int32_t h,
                           float* o_tmin, float* d_tmax);
/* Rays with origins uniform in the box [lo, hi] and directions uniform on
 * the sphere (tMin 0.001, tMax +inf); index i depends only on (seed, i). */
int32_t mt_gen_random_rays(uint64_t seed, int64_t first, int64_t n,
                           const float* lo, const float* hi,
                           float* o_tmin, float* d_tmax, int32_t nthreads);

/* UV spheres: `count` spheres of 64 slices x 33 rings (4096 triangles,
 * 2112 vertices with per-vertex normals each), radii lognormal(mu, sigma),
 * centres uniform in [c_lo, c_hi]^3, material index = sphere index, plus a
 * 2-triangle ground quad at y = ground_y spanning [-g, g]^2 with material
 * `count`.  Sizes: points/normals float4[2112*count + 4],
 * faces int32[10*(4096*count + 2)]. */
int32_t mt_gen_spheres(uint64_t seed, int32_t count, float mu, float sigma,
                       float c_lo, float c_hi, float ground_y, float g,
                       float* points, float* normals, int32_t* faces);

/* ------------------------------------------------------------------ */
/* Input formats (host)                                               */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t num_points, num_normals, num_uvs, num_tris, num_groups;
} mt_obj_counts;

/* OBJParser.readString (un-vendored coordinate.parser.obj.OBJParser, used
 * at TracerAPI.java:321-322): Wavefront OBJ text into the layouts above —
 * points/normals float4 (w = 0), uvs float2, faces int32[10] with -1 for an
 * absent uv/normal; polygons are fanned into triangles (corner 0, k, k+1);
 * negative indices count from the end.  Material index = rank of the
 * face's group among groups with faces (`o`, `g`, `usemtl` start one).
 * Call with points == NULL to get the counts, then with arrays of at least
 * those sizes in *counts (normals/uvs may be NULL). */
int32_t mt_obj_parse(const char* text, uint64_t len, mt_obj_counts* counts, float* points,
                     float* normals, float* uvs, int32_t* faces);
/* HDRBitmapReader (un-vendored bitmap.reader.HDRBitmapReader, used at
 * UserInterfaceFXMLController.java:176-179): Radiance RGBE (.hdr, "-Y h +X
 * w", flat, old-style or adaptive run-length scanlines) into float RGBA
 * rows (a = 1), value = mantissa * 2^(exponent - 136).  rgba == NULL
 * returns only the size. */
int32_t mt_hdr_decode(const uint8_t* data, uint64_t len, int32_t* width, int32_t* height, float* rgba);

/* ------------------------------------------------------------------ */
/* Device scene                                                       */
/* ------------------------------------------------------------------ */
typedef struct {
    const float* points;   /* device float4[num_points] */
    const float* normals;  /* device float4[], or NULL */
    const float* uvs;      /* device float2[], or NULL */
    const int32_t* faces;  /* device int32[10*num_tris] */
    const int32_t* nodes;  /* device int32[7*(2*num_tris-1)] */
    const float* bounds;   /* device float[8*(2*num_tris-1)] */
    int32_t num_tris;
    int32_t num_points;
    int32_t max_depth;     /* from mt_bvh_build; > 128 is rejected */
} mt_scene_desc;

/* CNormalBVH.build (CNormalBVH.java:45-148) on the device: from device
 * points/faces (layout above) into device nodes/bounds, the same tree as
 * mt_bvh_build (node numbering, leaf primitives, bounds; a zero bound may
 * differ in sign).  Builds one level of the tree per round and reads each
 * round's node count back, so the call synchronises the context's stream.
 * max_depth (host) receives the deepest leaf (root = 0). */
int32_t mt_bvh_build_device(mt_context* ctx, int32_t ntris, const float* points,
                            const int32_t* faces, int32_t* nodes, float* bounds,
                            int32_t* max_depth);

/* Validates the description and derives the packed traversal layout on the
 * device (the reference-layout arrays are left untouched and stay
 * readable).  CDeviceGI.set(mesh, bvh) + IntersectPrimitives' mesh/bvh
 * arguments (CDeviceGI.java:172, 405-411). */
int32_t mt_scene_create(mt_context* ctx, const mt_scene_desc* desc, mt_scene** out);
int32_t mt_scene_destroy(mt_scene* scene);

/* ------------------------------------------------------------------ */
/* A. Traversal                                                       */
/* ------------------------------------------------------------------ */
#define MT_TRACE_PACKED 0    /* default: packed 64-B node / 48-B triangle layout,
                                one ray per lane, short stack (= PACKED_STACK) */
#define MT_TRACE_REFERENCE 1 /* reads the reference's flat arrays directly */
#define MT_TRACE_PACKED_NO_STACK 2 /* packed, one ray per lane, parent-link backtracking */
#define MT_TRACE_PACKED_STACK 3    /* packed, one ray per lane, short stack */
#define MT_TRACE_PACKED_REFILL 4   /* packed, short stack, persistent waves with refill */
#define MT_TRACE_PACKED_SORTED 5   /* packed, short stack, each work-group's rays taken in
                                      direction-bin order (same results per ray) */
#define MT_TRACE_PACKED_POOL 6     /* packed, short stack, each wave refills finished lanes
                                      from its own run of 256 rays (same results per ray) */
#define MT_TRACE_PACKED_BINNED 7   /* packed: the batch counting-sorted by origin cell
                                      (32^3 grid, Morton order) on the device, then traced
                                      in that order by wave pools (same results per ray;
                                      for incoherent batches) */

/* Closest hit (intersectMesh with bailout = false, NormalBVH.cl:14-92, as
 * run by IntersectPrimitives :94-123; launched at CDeviceGI.java:230 and
 * CDeviceRT.java:176).  prim[i] = hit primitive or -1; t[i] = the ray's
 * final tMax (the hit distance on a hit).  Slots i >= *d_count are not
 * written.  stats, if non-NULL, receives per-slot {interior steps, leaf
 * visits, backtrack pops} as int32[3*capacity]. */
int32_t mt_trace_closest(mt_context* ctx, const mt_scene* scene,
                         const float* o_tmin, const float* d_tmax,
                         const int32_t* d_count, int32_t capacity,
                         int32_t* prim, float* t, int32_t* stats, int32_t mode);

/* Any hit (intersectOcclusion, NormalBVH.cl:319-348; testOcclusion
 * :235-317 inside DirectLight, Render.cl:456).  occluded[i] = 0/1. */
int32_t mt_trace_any(mt_context* ctx, const mt_scene* scene,
                     const float* o_tmin, const float* d_tmax,
                     const int32_t* d_count, int32_t capacity,
                     int32_t* occluded, int32_t* stats, int32_t mode);

/* Traversal autotuner (no counterpart in the reference, which has one
 * traversal kernel).  Every packed mode returns the same bits, but their
 * speed depends on the batch: one ray per lane with the short stack wins on
 * coherent rays, lane refill (persistent waves or wave pools) on incoherent
 * ones, ray binning on large incoherent ones.  Traces this batch (any = 0:
 * closest hit into prim/t; any != 0: any hit into occluded) with
 * MT_TRACE_PACKED_STACK, _REFILL, _POOL and _BINNED, `reps` timed runs each
 * after one warm-up, times them
 * with HIP events on the context stream and sets *best_mode to the fastest
 * and *best_ms to its mean time.  The outputs hold the batch's
 * result.  Synchronises the context. */
int32_t mt_trace_autotune(mt_context* ctx, const mt_scene* scene, const float* o_tmin,
                          const float* d_tmax, const int32_t* d_count, int32_t capacity,
                          int32_t any, int32_t* prim, float* t, int32_t* occluded, int32_t reps,
                          int32_t* best_mode, float* best_ms);

/* IntersectPrimitives' hit shading (NormalBVH.cl:124-232) for slots with
 * prim >= 0: p (after the shadow-terminator offset), shading normal n,
 * geometric normal ng (float4 each), uv (float2), material id (-1 on
 * miss). */
int32_t mt_shade_hits(mt_context* ctx, const mt_scene* scene,
                      const float* o_tmin, const float* d_tmax,
                      const int32_t* prim, const float* t,
                      const int32_t* d_count, int32_t capacity,
                      float* p, float* n, float* ng, float* uv, int32_t* mat);

/* IntersectPrimitives (NormalBVH.cl:94-233) over the reference's own
 * records, for a caller that holds its global Ray and Intersection buffers
 * (the kernel CDeviceRT.java:136 and CDeviceGI.java:172 create with rays,
 * isects, count, mesh and BVH arguments).  rays: `capacity` Ray records of
 * 80 bytes (Util.cl:129-138: float4 o, d, inv_d; int4 sign; int2 extra;
 * float tMin, tMax), read-write; isects: `capacity` Intersection records of
 * 96 bytes (Util.cl:141-151: float4 p, n, ng, d; float2 uv; int mat, id,
 * hit; 12 bytes padding).  Both 16-byte aligned.  For slots below *d_count
 * the closest hit is found in the reference's visiting order using the
 * record's inv_d and sign; a hit writes every Intersection field (p after
 * the shadow-terminator offset, w lanes 0; d copied from the ray; hit = 1)
 * and the hit distance into the ray's tMax.  A miss, and every slot in
 * [*d_count, capacity), writes only hit = 0 and mat = -1.  mode as for
 * mt_trace_closest (the packed modes share one kernel here). */
int32_t mt_intersect_primitives(mt_context* ctx, const mt_scene* scene, void* rays, void* isects,
                                const int32_t* d_count, int32_t capacity, int32_t mode);

/* intersectOcclusion (NormalBVH.cl:319-348) over the reference's 80-byte
 * Ray records (as mt_intersect_primitives): hits[i] = 1 if some triangle
 * meets the ray within (tMin, tMax) (intersectMesh with bailout), else 0,
 * for i < *d_count; later slots and the rays are left unchanged. */
int32_t mt_intersect_occlusion(mt_context* ctx, const mt_scene* scene, const void* rays, int32_t* hits,
                               const int32_t* d_count, int32_t capacity, int32_t mode);

/* Shadow rays toward a point light from each hit (Render.cl:453 with
 * Light.cl:93-97): origin getPoint(ray, t), direction (light - p)/dist,
 * tMin 0.0001, tMax dist - 2*EPS_RAY.  valid[i] = prim[i] >= 0. */
int32_t mt_shadow_rays(mt_context* ctx, const float* o_tmin, const float* d_tmax,
                       const int32_t* prim, const float* t, const float* light_xyz,
                       const int32_t* d_count, int32_t capacity,
                       float* s_o_tmin, float* s_d_tmax, int32_t* valid);

/* ------------------------------------------------------------------ */
/* D. Scan and compaction                                             */
/* ------------------------------------------------------------------ */

/* PrefixSumInteger.execute (PrefixSumInteger.java:80-88; kernels
 * Util.cl:948-980): out = exclusive scan of in[0..*d_len), d_total =
 * in[n-1] + out[n-1].  Single pass (decoupled look-back). */
int32_t mt_scan_excl_i32(mt_context* ctx, const int32_t* in, int32_t* out,
                         const int32_t* d_len, int32_t capacity, int32_t* d_total);

/* One payload stream moved by a compaction: element i of `src` (elem_bytes
 * 4, 8 or 16) goes to slot rank(i) of `dst`. */
typedef struct {
    const void* src;
    void* dst;
    int32_t elem_bytes;
} mt_payload;

/* Stable stream compaction of slots [0, *d_len) with flag != 0:
 * out_index[rank] = i (prepareLightInfo semantics, LightConfiguration.cl:
 * 31-55) and each payload moved to its rank; *d_count = survivors.
 * Replaces CCompact.execute's seven launches (CCompact.java:79-89) whose
 * CompactAtomic (Compact.cl:21-40) yields the same set in atomic order.
 * flag_mode 0: flags is int32 != 0; 1: flags is int32 >= 0 (a hit's
 * primitive id). out_index may be NULL.  npayload <= 8. */
int32_t mt_compact(mt_context* ctx, const int32_t* flags, int32_t flag_mode,
                   const int32_t* d_len, int32_t capacity, int32_t* out_index,
                   int32_t* d_count, const mt_payload* payloads, int32_t npayload);

/* mt_compact without allocation.  No counterpart in the reference: this is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as mt_compact, but the look-back
 * workspace is a fixed per-device array inside the library (256 KiB, held
 * by each device until the process ends) instead of the context's grown
 * workspace, so the call allocates nothing.  capacity must not exceed
 * MT_NOALLOC_MAX_ELEMS (MT_ERR_ARG otherwise).  All calls on one device
 * share that array: calls that overlap in time on one device (two
 * contexts, or two streams) race and give invalid results; calls ordered
 * on one stream are fine.  mt_scan_check reports its spin-limit flag. */
#define MT_NOALLOC_MAX_ELEMS (1 << 28)
int32_t mt_compact_noalloc(mt_context* ctx, const int32_t* flags, int32_t flag_mode,
                           const int32_t* d_len, int32_t capacity, int32_t* out_index,
                           int32_t* d_count, const mt_payload* payloads, int32_t npayload);

/* CCompact.execute (CCompact.java:79-89; CompactAtomic Compact.cl:21-40,
 * InitIntersection / InitIntData Initialize.cl:50-79, TransferIntersection /
 * TransferPixels Compact.cl:1-18) in place on the reference's own buffers:
 * of the first *d_len 96-byte Intersection records (Util.cl:141-151) those
 * with hit != 0 move to the front, each with its pixels[] entry; *d_count =
 * their number (CCompact's totalCount).  Records [*d_count, *d_len) become
 * InitIsect records (p, n, ng, d, uv 0; mat = id = -1; hit 0) with pixel 0;
 * padding bytes are written 0; slots from *d_len on are untouched.
 * Survivors keep index order (CompactAtomic's order is that of its atomic
 * increments; the set and count are the same).  temp_isects: caller-owned
 * scratch of `capacity` records (CCompact's tempIsectBuffer); no temp pixel
 * buffer is needed.  Records 16-byte aligned. */
int32_t mt_compact_isects(mt_context* ctx, void* isects, int32_t* pixels, void* temp_isects,
                          const int32_t* d_len, int32_t capacity, int32_t* d_count);

/* Synchronises the context and reports MT_ERR_STATE if a scan or compaction
 * since the last check gave up waiting on a predecessor tile (a bounded spin
 * that should never trigger); clears the flag.  No counterpart in the
 * reference, whose scans have no error channel. */
int32_t mt_scan_check(mt_context* ctx);

/* identifyMeshLights + PrefixSumInteger + prepareLightInfo
 * (CLightConfiguration.initLights, CLightConfiguration.java:43-87):
 * light_faces[k] = k-th face (in index order) whose material is an emitter
 * (Luminance(emission_color * emission_param.x) != 0, Material.cl:27-31). */
int32_t mt_mesh_lights(mt_context* ctx, const mt_scene* scene, const float* materials,
                       int32_t* light_faces, int32_t* d_count);

/* ------------------------------------------------------------------ */
/* C. Adaptive light grid and environment SATs                        */
/* ------------------------------------------------------------------ */

/* LightGrid (LightGrid.cl:9-37): nu = nv = 1600, tiles 16 x 32, 100 x 50
 * cells; accum/func/sat are float[2 560 000] each, row-major 1600 wide. */
int32_t mt_lightgrid_create(mt_context* ctx, mt_lightgrid** out);
int32_t mt_lightgrid_destroy(mt_lightgrid* grid);
/* Raw device arrays.  Sampling reads per-cell tables (the sampler's SAT
 * quotients, computed once per SAT build); asking for func or sat switches
 * sampling back to the SAT itself until the next clear / set_tile_func /
 * update / build_sat, so direct writes through these pointers are honoured. */
int32_t mt_lightgrid_arrays(mt_lightgrid* grid, float** accum, float** func, float** sat);
/* initLightGrid (LightGrid.cl:404-410; CAdaptiveEnvironment.clearBuffers :123-127) */
int32_t mt_lightgrid_clear(mt_context* ctx, mt_lightgrid* grid);
/* setToAllFunc + row/col SAT (CAdaptiveEnvironment.setHDRLuminance :69-88):
 * copies the 512-entry tile `tile` (host, x fastest) into every cell, then
 * builds the SAT. */
int32_t mt_lightgrid_set_tile_func(mt_context* ctx, mt_lightgrid* grid, const float* tile);
/* CAdaptiveEnvironment.update (:135-145): func <- accum, then the tile SAT
 * (initFuncAndSat, calculateFunc, prefixSumRowSubgrid, prefixSumColSubgrid
 * in one kernel). */
int32_t mt_lightgrid_update(mt_context* ctx, mt_lightgrid* grid);
/* prefixSumRowSubgrid + prefixSumColSubgrid on the current func. */
int32_t mt_lightgrid_build_sat(mt_context* ctx, mt_lightgrid* grid);
/* sampleSubgridContinuous (LightGrid.cl:310-320) for n (cell, u0, u1). */
int32_t mt_lightgrid_sample(mt_context* ctx, const mt_lightgrid* grid, int32_t n,
                            const int32_t* cell, const float* u, int32_t* tile_xy,
                            float* uv, float* pdf);
/* subgridIndexFromCamera / tileIndexXYFromCamera (LightGrid.cl:68-90). */
int32_t mt_lightgrid_cells(mt_context* ctx, int32_t n, const float* dirs,
                           int32_t* cell, int32_t* tile_xy);
/* Generic tile SAT over a nu x nv grid with rx-wide row and ry-tall column
 * segments (the kernels above on other grid sizes, e.g. the 16x16 grid of
 * test/lightgrid/SimpleGrid.cl). rx, ry powers of two <= 64. */
int32_t mt_sat_tiles(mt_context* ctx, const float* func, float* sat,
                     int32_t nu, int32_t nv, int32_t rx, int32_t ry);

/* Environment map: rgba float4[w*h]; lum = Luminance(rgb)
 * (InitEnvironmentLum, Initialize.cl:100-115); sat = calculateSAT order
 * (SAT.cl:304-310: sequential rows, then sequential columns). */
int32_t mt_env_lum_sat(mt_context* ctx, const float* rgba, int32_t w, int32_t h,
                       float* lum, float* sat);
/* Holds device pointers to rgba/lum/sat (caller-owned) for the light calls. */
int32_t mt_env_create(mt_context* ctx, const float* rgba, const float* lum,
                      const float* sat, int32_t w, int32_t h, mt_envmap** out);
int32_t mt_env_destroy(mt_envmap* env);
/* sampleContinuous on a sub-region (SAT.cl:244-323). */
int32_t mt_sat_sample_region(mt_context* ctx, const float* func, const float* sat,
                             int32_t nu, int32_t nv, int32_t minx, int32_t miny,
                             int32_t sx, int32_t sy, int32_t n, const float* u,
                             float* uv, int32_t* offset, float* pdf);
/* illuminateEnvironmentLight (Light.cl:195-227). */
int32_t mt_env_illuminate(mt_context* ctx, const mt_envmap* env, int32_t n,
                          const float* u, float* dir, float* contrib, float* pdfw,
                          int32_t* texel);
/* getRadianceEnvironmentLight (Light.cl:270-286). */
int32_t mt_env_radiance(mt_context* ctx, const mt_envmap* env, int32_t n,
                        const float* dir, float* contrib, float* pdfw, int32_t* texel);

/* CAdaptiveEnvironment.setCameraPosition (CAdaptiveEnvironment.java:61-67,
 * called per frame by CDeviceGI.java:296): pos = host float[3]; the grid's
 * cells are spherical bins of the direction camera -> hit point.  Taken by
 * the adaptive calls below when they are enqueued.  Initially (0, 0, 0). */
int32_t mt_lightgrid_set_camera(mt_lightgrid* grid, const float* pos);
/* illuminateEnvironmentLight1 (Light.cl:140-193), the light-grid-guided
 * environment sample (no live caller in the reference: Render.cl:458-459
 * are commented out).  For each of n receiving points hit_p (float4): cell
 * from the camera direction, tile ~ the cell's distribution with (u.x, u.y)
 * (sampleSubgridContinuous), texel ~ that tile's env-map SAT region with
 * (u.z, u.w) (u: float4[n]).  Outputs dir (float4, w 0), contrib (env
 * texel), pdfw (the in-tile solid-angle pdf, as the reference: the tile
 * pick's pdf is not applied), and, if non-NULL, lum = Luminance(contrib)
 * and grid_index = the accum entry of (cell, tile).  The env map needs
 * w >= 16 and h >= 32 (one texel per tile). */
int32_t mt_env_illuminate_adaptive(mt_context* ctx, const mt_envmap* env, const mt_lightgrid* grid,
                                   int32_t n, const float* hit_p, const float* u, float* dir,
                                   float* contrib, float* pdfw, float* lum, int32_t* grid_index);
/* getRadianceEnvironmentLight1 (Light.cl:229-268) for n escaped rays
 * (dir float4, origin hit_p float4): tile from the ray direction, pdfa =
 * getPdfSAT over that tile's region / (2 pi^2 sin(pi y/h)), contrib = env
 * texel.  accumulate != 0 adds Luminance(contrib) into the grid's accum
 * entry of (cell, tile) (accumLightGrid, Light.cl:35-38, float atomic add:
 * per-entry sums are order-dependent in their last bits, as the
 * reference's).  contrib, pdfa, lum, grid_index may be NULL. */
int32_t mt_env_radiance_adaptive(mt_context* ctx, const mt_envmap* env, mt_lightgrid* grid, int32_t n,
                                 const float* dir, const float* hit_p, int32_t accumulate,
                                 float* contrib, float* pdfa, float* lum, int32_t* grid_index);
/* The environment's light grid (EnvironmentLight.lightGrid, Light.cl:15-33;
 * set by getEnvironmentLight, :129-137), used by renders with
 * MT_RENDER_LIGHTGRID_ENV; NULL detaches.  The grid must outlive its use. */
int32_t mt_env_attach_lightgrid(mt_envmap* env, mt_lightgrid* grid);

/* ------------------------------------------------------------------ */
/* B. BSDF                                                            */
/* ------------------------------------------------------------------ */
#define MT_BSDF_WORDS 20 /* frame mX,mY,mZ (9), localDirFix (3), localGeomN (3),
                            diffProb, glossyProb, 3 reserved */
/* setupBsdf (Bsdf.cl:130-160) for n hits: ray_d, isect_n, isect_ng are
 * float4[n] (xyz used), mat int32[n] (material index, or -1 for a miss);
 * bsdf float[20*n], mat_id (-1 for a miss or a grazing incoming
 * direction). */
int32_t mt_bsdf_setup(mt_context* ctx, int32_t n, const float* ray_d,
                      const float* isect_n, const float* isect_ng, const int32_t* mat,
                      const float* materials, float* bsdf, int32_t* mat_id);
/* SampleBsdf (Bsdf.cl:555-606): u float[3n]; f, dir float[3n]; pdf and
 * cos float[n]. */
int32_t mt_bsdf_sample(mt_context* ctx, int32_t n, const float* bsdf, const int32_t* mat_id,
                       const float* materials, const float* u, float* f, float* dir,
                       float* pdf, float* cos_out);
/* EvaluateBsdf (Bsdf.cl:525-545): dir float[3n] world directions. */
int32_t mt_bsdf_eval(mt_context* ctx, int32_t n, const float* bsdf, const int32_t* mat_id,
                     const float* materials, const float* dir, float* f, float* pdf,
                     float* cos_out);

/* ------------------------------------------------------------------ */
/* Wavefront bounce loop                                              */
/* ------------------------------------------------------------------ */
typedef struct {
    float pos[3];          /* pinhole camera position */
    float look[3];         /* look-at point */
    float up[3];
    float vfov_deg;
    int32_t width, height; /* image size */
    int32_t spp;           /* samples per pixel, all traced as one wavefront */
    int32_t max_bounces;   /* path segments traced per sample (the reference fixes 15) */
    int32_t rr_start;      /* Russian roulette from this bounce on; < 0 disables */
    uint32_t seed;         /* job seed: RNG keys are (seed, path id, bounce, purpose) */
    int32_t mode;          /* MT_TRACE_PACKED or MT_TRACE_REFERENCE traversal */
    int32_t flags;         /* MT_RENDER_* bits */
} mt_render_desc;

/* illuminateMeshLight reads lightInfo->faceId, i.e. always the FIRST light
 * (Render.cl:304-305), while dividing the pick probability by the light
 * count; the default reproduces that, this flag samples the picked light. */
#define MT_RENDER_PICKED_LIGHT 1
/* time every traversal launch with HIP events (stats->closest_ms/any_ms) */
#define MT_RENDER_PROFILE 2
/* run the counting traversal variant and total its per-ray counters */
#define MT_RENDER_TRAVERSAL_STATS 4
/* Environment light guided by the light grid attached to env
 * (mt_env_attach_lightgrid): DirectLight samples it with
 * illuminateEnvironmentLight1 (Light.cl:140-193; the grid cell of the
 * receiving point, a tile from the cell's distribution, a texel from the
 * tile's SAT region) and escaped rays take getRadianceEnvironmentLight1
 * (:229-268) at the path's last hit point ((0,0,0) for camera rays,
 * InitPathData, Initialize.cl:46), the pairing Render.cl:458-459 leaves
 * commented out.  Every escaped ray, and every unoccluded environment light
 * sample (Render.cl:458-459), adds the texel's luminance to the grid entry
 * of (cell, tile) (accumLightGrid, Light.cl:35-38): in fixed point (2^-20)
 * with integer atomics during the render, so the sums do not depend on the
 * order of the additions, then added once per entry to the grid's float
 * accum at the end of the render.  The grid's camera position and sampling
 * state are read when the render is enqueued (for a captured render: at
 * mt_render_graph_create; the grid arrays keep their addresses, so updates
 * between replays are seen).  Callers update the grid between frames as
 * CDeviceGI.loop does for frames 0-4 (mt_lightgrid_update,
 * CDeviceGI.java:263-264). */
#define MT_RENDER_LIGHTGRID_ENV 8
/* With MT_RENDER_LIGHTGRID_ENV: the solid-angle pdf of a guided sample, and
 * of an escaped ray's direction, includes the probability density of the
 * tile pick (the reference computes it as pdfLL and leaves it unused,
 * Light.cl:163,182, so its estimator is biased wherever the cell's tile
 * distribution is not uniform).  Without this flag the reference's in-tile
 * pdf is used.  Exact for maps whose width is a multiple of 16 and height a
 * multiple of 32 (texel rectangles of equal size). */
#define MT_RENDER_LIGHTGRID_TILE_PDF 16
/* MT_RENDER_TAIL_FROM(b), 1 <= b < max_bounces (flags bits 8-13): bounces b
 * and later run in one launch in which each lane carries a path through all
 * of its remaining bounces (closest hit, shading, shadow ray) instead of
 * three launches per bounce.  Same operations per path in the same order,
 * so images, ray counts and light-grid sums are bit-identical; it pays when
 * few paths remain at bounce b (open scenes, where most paths escape early:
 * each late bounce's launches otherwise last as long as their slowest ray),
 * and costs when many do (the fused wavefront shares work across paths
 * better).  Applies to the default packed traversal modes without
 * MT_RENDER_TRAVERSAL_STATS; 0 (the default) or b >= max_bounces: off.  No
 * counterpart in the reference (its loop, CDeviceGI.java:214-271, launches
 * every stage of every bounce). */
#define MT_RENDER_TAIL_SHIFT 8
#define MT_RENDER_TAIL_FROM(b) ((int32_t)(b) << MT_RENDER_TAIL_SHIFT)

typedef struct {
    int64_t extension_rays; /* closest-hit rays traced (all bounces) */
    int64_t shadow_rays;    /* any-hit rays traced */
    int64_t paths;          /* width * height * spp */
    int32_t bounce_counts[64]; /* live paths entering each bounce */
    /* MT_RENDER_PROFILE: summed durations of the traversal launches, timed
     * with HIP events on the context stream */
    int32_t closest_launches, any_launches;
    float closest_ms, any_ms;
    /* MT_RENDER_TRAVERSAL_STATS: totals of {interior steps, leaf visits,
     * backtrack pops} over all closest-hit / any-hit rays */
    int64_t closest_ilp[3], any_ilp[3];
} mt_render_stats;

/* CDeviceGI.loop (src/cl/device/CDeviceGI.java:214-271) for one job:
 * camera rays, then per bounce IntersectPrimitives, SetupBSDFPath,
 * LightHitPass (environment miss radiance with MIS, emitter hits),
 * DirectLight (environment sample through the full-map SAT, BSDF
 * evaluation, any-hit shadow ray), SampleBSDFRayDirection, and the live
 * paths gathered into the next bounce's queue (the default traversal
 * appends them while shading; with MT_RENDER_TRAVERSAL_STATS or the
 * reference layout a stable compaction does it; the image is the same
 * either way).  materials is a device array of 36-word records;
 * env may be NULL (no environment light); light_faces/num_lights (device,
 * from mt_mesh_lights) may be NULL/0.  image receives float4[w*h]: the mean
 * over spp of each pixel's radiance.  RNG is keyed by path id, so a path's
 * result does not depend on which other paths share the wavefront.
 * The wavefront state (about 100 B per path) lives in the context's
 * grow-only workspace, so repeated renders allocate nothing; it is held
 * until mt_ctx_destroy.  Synchronises the context before returning.
 * stats (host) may be NULL. */
int32_t mt_render(mt_context* ctx, const mt_scene* scene, const float* materials,
                  const mt_envmap* env, const int32_t* light_faces,
                  const int32_t* d_num_lights, const mt_render_desc* desc,
                  float* image, mt_render_stats* stats);

/* Captured render.  The reference's render thread re-issues the same
 * launch sequence for every frame (CDeviceGI.loop, CDeviceGI.java:214-271,
 * once per frame from execute(), :424-431).  mt_render_graph_create lays out
 * one job's wavefront state once (library-owned, freed by destroy) and
 * records mt_render's whole launch sequence into a HIP graph;
 * mt_render_graph_launch(g, seed) replays it on the context stream as one
 * launch, with `seed` in place of desc->seed, and produces exactly the
 * image mt_render produces for that seed.  The scene, materials,
 * environment, light list and image must outlive the graph and keep their
 * addresses.  desc->flags may not hold MT_RENDER_PROFILE.
 * mt_render_graph_stats fills the ray counts (and, with
 * MT_RENDER_TRAVERSAL_STATS, the traversal totals) of the last replay;
 * it synchronises the context.  Replays of one graph are serialised on the
 * context stream; one graph must not be replayed on two streams at once. */
typedef struct mt_render_graph mt_render_graph;
int32_t mt_render_graph_create(mt_context* ctx, const mt_scene* scene, const float* materials,
                               const mt_envmap* env, const int32_t* light_faces,
                               const int32_t* d_num_lights, const mt_render_desc* desc,
                               float* image, mt_render_graph** out);
int32_t mt_render_graph_launch(mt_render_graph* g, uint32_t seed);
int32_t mt_render_graph_stats(mt_render_graph* g, mt_render_stats* stats);
int32_t mt_render_graph_destroy(mt_render_graph* g);

/* ------------------------------------------------------------------ */
/* Image pipeline                                                     */
/* ------------------------------------------------------------------ */
/* CImage.initBuffers (CImage.java:117-122; InitFloat4DataXYZ and
 * InitIntDataRGB, Initialize.cl:1-11,91-98): accum[i] = (0,0,0,1) and
 * argb[i] = 0xFF000000; either pointer may be NULL. */
int32_t mt_image_init(mt_context* ctx, float* accum, int32_t* argb, int32_t n);
/* CImage.processImage (CImage.java:124-135): averageAccum (Image.cl:29-65)
 * frame = accum / frame_count and per-pixel log(0.01 + Lw); when
 * update_totals, *d_total_log / *d_total_count = sum of the logs / count of
 * pixels with Lw > 0 (fixed-order float sum; the reference recomputes them
 * while frameCount < 3); then updateFrameImage (Image.cl:68-118): Reinhard
 * tone map of the luminance scaled by exposure / exp(mean log), xyY -> RGB,
 * gamma, argb = getIntARGB.  Pixels with Lw <= 0 keep their argb value.
 * frame: float4[n] out; argb: int32[n]; totals: device scalars. */
int32_t mt_image_process(mt_context* ctx, const float* accum, float frame_count, int32_t n,
                         float exposure, float gamma, int32_t update_totals, float* frame,
                         int32_t* argb, float* d_total_log, int32_t* d_total_count);
/* PrefixSumFloat.execute (PrefixSumFloat.java:95-102; kernels
 * Util.cl:982-1042): out = exclusive scan of in[0..*d_len), *d_total = the
 * sum (d_total may be NULL).  The reference's global pass converts partial
 * sums to int (Util.cl:1035); this scan is a float scan in a fixed order
 * (deterministic run to run). */
int32_t mt_scan_excl_f32(mt_context* ctx, const float* in, float* out, const int32_t* d_len,
                         int32_t capacity, float* d_total);

/* ------------------------------------------------------------------ */
/* Multi-job wavefront engine (continuous batching)                   */
/* ------------------------------------------------------------------ */
/* The reference renders one job per device at a time: CDeviceGI.start
 * hands execute() -> loop() to its render thread (CDeviceGI.java:190-204,
 * 424-431) and every launch covers that job's paths only.  mt_batch runs
 * many jobs through ONE set of wavefront queues: each iteration refills the
 * free queue slots with new camera paths of the admitted jobs (oldest job
 * first), traces all live rays against their own job's scene in one
 * closest-hit and one any-hit launch, shades every slot with its job's
 * materials and lights, and compacts the survivors.  A path computes the
 * same bits as under mt_render (same per-path code, RNG keyed by job seed
 * and path id), so every job's image is identical to a plain mt_render of
 * that job, alone or mixed with any others.
 *
 * Jobs are admitted in submission order when a job slot is free and their
 * per-path radiance buffer (mt_batch_job_bytes) fits in the engine's arena;
 * scenes, materials, environments and images stay owned by the caller and
 * must outlive the job. */
typedef struct mt_batch mt_batch;

typedef struct {
    const mt_scene* scene;        /* as mt_render */
    const float* materials;
    const mt_envmap* env;
    const int32_t* light_faces;
    const int32_t* d_num_lights;
    mt_render_desc render;        /* mode must be MT_TRACE_PACKED or
                                     MT_TRACE_PACKED_STACK; flags may hold
                                     MT_RENDER_PICKED_LIGHT only */
    float* image;                 /* device float4[w*h], written at completion */
    int64_t tag;                  /* caller's id, returned by mt_batch_step */
} mt_job_desc;

typedef struct {
    int64_t iterations;           /* wavefront iterations run */
    int64_t extension_rays;       /* closest-hit rays traced */
    int64_t shadow_rays;          /* any-hit rays with a light sample */
    int64_t jobs_submitted, jobs_admitted, jobs_completed, jobs_cancelled;
    int32_t active_jobs, pending_jobs;
    uint64_t arena_used, arena_peak; /* bytes of per-job buffers held */
} mt_batch_stats;

/* queue_capacity: live paths per iteration (1 .. 2^30); arena_bytes: device
 * memory for per-job radiance buffers; max_jobs: jobs admitted at once
 * (1 .. 1<<20). */
int32_t mt_batch_create(mt_context* ctx, int32_t queue_capacity, uint64_t arena_bytes, int32_t max_jobs,
                        mt_batch** out);
int32_t mt_batch_destroy(mt_batch* batch);
/* Arena bytes a job holds while admitted (16 B per path, 256-B aligned). */
int32_t mt_batch_job_bytes(const mt_render_desc* desc, uint64_t* bytes);
/* Queues a job (validated now; MT_ERR_CAPACITY if it can never fit). */
int32_t mt_batch_submit(mt_batch* batch, const mt_job_desc* job);
/* Admits what fits, runs the wavefront iterations of one step (1 by
 * default, mt_batch_set_iterations) and synchronises.  Writes the tags of
 * jobs whose image is complete, or whose cancellation has drained, (at most
 * max_done; the rest follow on later calls) to done_tags, their number to
 * *n_done, and *busy = 1 while any job is pending or in flight. */
int32_t mt_batch_step(mt_batch* batch, int32_t max_done, int64_t* done_tags, int32_t* n_done, int32_t* busy);
int32_t mt_batch_get_stats(mt_batch* batch, mt_batch_stats* stats);
/* Which admitted jobs get free queue slots first: the oldest (default), or
 * the one with the fewest paths (short jobs finish within max_bounces
 * iterations of admission, ahead of long ones).  Images do not depend on it. */
#define MT_BATCH_OLDEST_FIRST 0
#define MT_BATCH_SMALLEST_FIRST 1
int32_t mt_batch_set_policy(mt_batch* batch, int32_t policy);
/* Wavefront iterations per mt_batch_step (1 .. 1024, default 1): more
 * iterations between host round trips; a job that finishes early in a step
 * is resolved and reported at the end of the step. */
int32_t mt_batch_set_iterations(mt_batch* batch, int32_t iterations);
/* No counterpart in the reference (one job per device, one launch per
 * kernel): on = 1 (default) records a step's iterations once as a hipGraph
 * per start parity of the queue ping-pong and replays it every step (one
 * graph launch instead of about ten kernel launches per iteration; every
 * size the iterations use lives in device memory, so the recording stays
 * valid while jobs come and go, until the iterations per step change);
 * on = 0 launches the kernels one by one.  Images do not depend on it.
 * Synchronises the context. */
int32_t mt_batch_set_graphs(mt_batch* batch, int32_t on);
/* Closest-hit traversal of the engine's shared queue (no counterpart in the
 * reference): MT_BATCH_TRACE_SPLIT (MT_TRACE_PACKED, the default: the
 * queue's front, the paths continuing from earlier bounces, with wave pools;
 * its back, the camera rays of newly started paths, one ray per lane),
 * MT_TRACE_PACKED_STACK (one ray per lane) or MT_TRACE_PACKED_POOL (wave
 * pools with lane refill; each lane loads its slot's scene).  Images do not
 * depend on it.  The choice the serving autotuner makes per workload
 * (multi-job workloads of 256 uneven jobs, and of Poisson arrivals: split
 * 64.0 jobs/s, one ray per lane 60.4, pools 59.8).  Synchronises the
 * context. */
#define MT_BATCH_TRACE_SPLIT 8
int32_t mt_batch_set_trace_mode(mt_batch* batch, int32_t mode);
/* Cancels the job with this tag (CDeviceGI.stop, CDeviceGI.java:438-439,
 * ends the reference's one running job; here any one job of the mix).
 * A job not yet admitted is dropped at once; an admitted job is issued no
 * more paths and its paths in the queues end at their next shading, so it
 * drains within one iteration.  Either way its tag is reported by a later
 * mt_batch_step like a completed job's, its image is not written, and its
 * slot and arena space are then free.  Other jobs' images are unaffected
 * (bit-identical to their plain renders).  *found = 0 when no pending or
 * running job has the tag. */
int32_t mt_batch_cancel(mt_batch* batch, int64_t tag, int32_t* found);

#ifdef __cplusplus
}
#endif
#endif
