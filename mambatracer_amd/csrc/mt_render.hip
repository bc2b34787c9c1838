// Wavefront bounce loop: restates CDeviceGI.loop (src/cl/device/CDeviceGI.java:
// 214-271) and its kernels SetupBSDFPath, LightHitPass, DirectLight and
// SampleBSDFRayDirection (src/cl/kernel/Render.cl:10-125,233-466).
//
// Layout.  Path state is SoA indexed by path id (pid = sample * pixels +
// pixel): throughput.xyz + last BSDF pdf in one float4, accumulated
// radiance in another.  The live paths of a bounce form a queue of pids with
// their rays (two float4 streams) and a device-resident count; every stage
// reads the count on the device (the reference's `global int* count`), so
// the host enqueues the whole render without synchronising.
//
// Per bounce:
//   1. closest hit over the queue (mt_trace_closest);
//   2. one fused shading kernel per slot: IntersectPrimitives' hit shading,
//      setupBsdf, LightHitPass (environment radiance on a miss with MIS,
//      emitter hit), DirectLight's light sample + EvaluateBsdf + shadow-ray
//      setup, SampleBsdf for the next ray and Russian roulette;
//   3. any hit over the shadow rays (mt_trace_any);
//   4. unoccluded direct light added to the path radiance;
//   5. stable compaction of live slots (pid and next ray) into the other
//      queue buffer (mt_compact; CCompact.execute in the reference).
// After the last bounce the per-path radiance is averaged per pixel in a
// fixed order, so the image is deterministic: no float atomics.
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
#include "mt_bsdf.h"
#include "mt_sat.h"

namespace {

constexpr int RT = 256; // threads per block

struct Camera {
    float ox, oy, oz;
    float fx, fy, fz, rx, ry, rz, ux, uy, uz;
    float tan_half, aspect;
    int w, h;
};

struct PathState {
    float4* T;   // throughput.xyz, last BSDF pdf (lastPdfW)
    float4* L;   // accumulated radiance
};

struct SceneView {
    const float4* points;
    const float4* normals;
    const int* faces;
    const MtMaterial* mats;
    const int* light_faces;
    const int* num_lights;
    EnvView env;
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

__global__ __launch_bounds__(RT) void k_camera(Camera C, int npix, int npaths, uint32_t seed,
                                               float4* __restrict__ o_tmin, float4* __restrict__ d_tmax,
                                               int* __restrict__ queue, PathState S, int* __restrict__ count0)
{
    int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid == 0) *count0 = npaths;
    if (pid >= npaths) return;
    int pixel = pid % npix;
    int px = pixel % C.w, py = pixel / C.w;
    int32_t sx, sy;
    seed_for(seed, pid, 0, 0, sx, sy);
    float j0 = mt_get_random(sx, sy);
    float j1 = mt_get_random(sx, sy);
    float a = (2.0f * ((float)px + j0) / (float)C.w - 1.0f) * C.tan_half * C.aspect;
    float b = (1.0f - 2.0f * ((float)py + j1) / (float)C.h) * C.tan_half;
    v3 d = norm3(mk3(C.fx + a * C.rx + b * C.ux, C.fy + a * C.ry + b * C.uy, C.fz + a * C.rz + b * C.uz));
    o_tmin[pid] = make_float4(C.ox, C.oy, C.oz, 0.001f);   // initGlobalRay, Util.cl:577-587
    d_tmax[pid] = make_float4(d.x, d.y, d.z, INFINITY);
    queue[pid] = pid;
    S.T[pid] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);         // InitPathData, Initialize.cl:25-48
    S.L[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
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

struct ShadeArgs {
    SceneView V;
    PathState S;
    const int* queue;
    const int* d_count;
    float4* o_tmin;       // in: this bounce's rays; out: next rays (same slot)
    float4* d_tmax;
    const int* prim;
    const float* tin;
    int* alive;           // out: 1 if the slot's path continues
    float4* s_o_tmin;     // out: shadow rays (tMin = inf marks "none")
    float4* s_d_tmax;
    float4* s_contrib;    // out: unoccluded direct-light contribution
    int* shadow_count;
    int bounce;
    int rr_start;
    int flags;
    uint32_t seed;
};

// One slot of the fused shading stage; returns whether a shadow ray was set.
__device__ __forceinline__ bool shade_slot(const ShadeArgs& A, int slot)
{
    bool has_shadow = false;
    {
        const SceneView& V = A.V;
        int pid = A.queue[slot];
        float4 ro = A.o_tmin[slot], rd = A.d_tmax[slot];
        v3 d = xyz(rd);
        float4 T4 = A.S.T[pid];
        v3 T = xyz(T4);
        float last_pdf = T4.w;
        bool last_specular = A.bounce == 0; // InitPathData sets it, SampleBSDFRayDirection clears it
        v3 Ladd = mk3(0, 0, 0);
        int c = A.prim[slot];
        bool alive = false;
        float4 so = make_float4(0, 0, 0, INFINITY), sd = make_float4(1, 0, 0, -1.0f);
        float4 scontrib = make_float4(0, 0, 0, 0);
        if (c < 0) {
            // LightHitPass, environment branch (Render.cl:106-123)
            if (V.env.w > 0) {
                float pdfw;
                int tx;
                float4 cv = env_radiance(V.env, d, pdfw, tx);
                v3 contrib = xyz(cv);
                if (!(contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f)) {
                    float w = last_specular ? 1.0f : mis2(last_pdf, pdfw);
                    v3 v = mk3(T.x * contrib.x * w, T.y * contrib.y * w, T.z * contrib.z * w);
                    if (rgb_valid(v)) Ladd = v;
                }
            }
        } else {
            // IntersectPrimitives hit shading (NormalBVH.cl:124-220)
            const int* F = V.faces + 10 * c;
            v3 p1 = xyz(V.points[F[0]]), p2 = xyz(V.points[F[1]]), p3 = xyz(V.points[F[2]]);
            float tt = A.tin[slot];
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
                    if (cos_out != 0.0f && !(AL.intensity.x == 0.0f && AL.intensity.y == 0.0f &&
                                             AL.intensity.z == 0.0f)) {
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
                    seed_for(A.seed, pid, A.bounce, 1, sx, sy);
                    float s0 = mt_get_random(sx, sy), s1 = mt_get_random(sx, sy);
                    float s2 = mt_get_random(sx, sy), s3 = mt_get_random(sx, sy);
                    (void)s2;
                    (void)s3;
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
                    if (use_env) { // illuminateEnvironmentLight, Light.cl:195-227
                        int tx;
                        float4 cv = env_illuminate(V.env, s0, s1, ldir, lpdf, tx);
                        rad = xyz(cv);
                        dist = MT_FLOATMAX;
                    } else if (use_mesh) { // illuminateMeshLight, Render.cl:287-315
                        int li = (int)(mt_get_random(sx, sy) * (float)nlights);
                        int tri = (A.flags & MT_RENDER_PICKED_LIGHT) ? V.light_faces[li] : V.light_faces[0];
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
                                so = make_float4(p.x, p.y, p.z, 0.0001f);
                                sd = make_float4(ldir.x, ldir.y, ldir.z, dist - 2 * MT_EPS_RAY);
                                scontrib = make_float4(cv.x, cv.y, cv.z, 1.0f);
                                has_shadow = true;
                            }
                        }
                    }
                    // SampleBSDFRayDirection (Render.cl:233-285)
                    int32_t bx, by;
                    seed_for(A.seed, pid, A.bounce, 2, bx, by);
                    float u0 = mt_get_random(bx, by), u1 = mt_get_random(bx, by), u2 = mt_get_random(bx, by);
                    v3 wdir;
                    float pdf = 0.0f, cosg = 0.0f;
                    v3 f = mt_bsdf_sample_one(B, M, u0, u1, u2, wdir, pdf, cosg);
                    float k = cosg / pdf;
                    v3 Tn = mk3(T.x * (f.x * k), T.y * (f.y * k), T.z * (f.z * k));
                    alive = pdf > 0.0f && !(f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) && rgb_valid(Tn) &&
                            !(Tn.x == 0.0f && Tn.y == 0.0f && Tn.z == 0.0f);
                    if (alive && A.rr_start >= 0 && A.bounce + 1 >= A.rr_start) {
                        float q = fminf(1.0f, fmaxf(Tn.x, fmaxf(Tn.y, Tn.z)));
                        float r = mt_get_random(bx, by);
                        if (r >= q) alive = false;
                        else Tn = mk3(Tn.x / q, Tn.y / q, Tn.z / q);
                    }
                    if (alive) {
                        A.S.T[pid] = make_float4(Tn.x, Tn.y, Tn.z, pdf);
                        A.o_tmin[slot] = make_float4(p.x, p.y, p.z, 0.001f); // initGlobalRay
                        A.d_tmax[slot] = make_float4(wdir.x, wdir.y, wdir.z, INFINITY);
                    }
                }
            }
        }
        if (Ladd.x != 0.0f || Ladd.y != 0.0f || Ladd.z != 0.0f) {
            float4 L = A.S.L[pid];
            A.S.L[pid] = make_float4(L.x + Ladd.x, L.y + Ladd.y, L.z + Ladd.z, L.w);
        }
        A.alive[slot] = alive ? 1 : 0;
        A.s_o_tmin[slot] = so;
        A.s_d_tmax[slot] = sd;
        A.s_contrib[slot] = scontrib;
    }
    return has_shadow;
}

// Persistent: waves stride over the live slots [0, *d_count).
__global__ __launch_bounds__(RT) void k_shade(ShadeArgs A)
{
    const int count = *A.d_count;
    const int lane = __lane_id();
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int base = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64; base < count; base += nw * 64) {
        int slot = base + lane;
        bool sh = slot < count ? shade_slot(A, slot) : false;
        uint64_t m = __ballot(sh);
        if (m && lane == 0) atomicAdd(A.shadow_count, (int)__popcll(m));
    }
}

__global__ __launch_bounds__(RT) void k_nee_accum(const int* __restrict__ queue, const int* __restrict__ d_count,
                                                  const float4* __restrict__ contrib,
                                                  const int* __restrict__ occluded, PathState S)
{
    const int count = *d_count;
    for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < count; slot += gridDim.x * blockDim.x) {
        float4 c = contrib[slot];
        if (c.w == 0.0f || occluded[slot]) continue;
        int pid = queue[slot];
        float4 L = S.L[pid];
        S.L[pid] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
    }
}

__global__ __launch_bounds__(RT) void k_resolve(const float4* __restrict__ L, int npix, int spp,
                                                float4* __restrict__ image)
{ // mean over samples in sample order (averageAccum, Image.cl:29-45, without tone mapping)
    int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= npix) return;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    for (int s = 0; s < spp; s++) {
        float4 v = L[(size_t)s * npix + px];
        r += v.x;
        g += v.y;
        b += v.z;
    }
    float inv = 1.0f / (float)spp;
    image[px] = make_float4(r * inv, g * inv, b * inv, 1.0f);
}

__global__ void k_zero_i32(int* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Sums per-slot traversal counters {I, L, P} of the slots < *d_count.
__global__ __launch_bounds__(RT) void k_sum_ilp(const int* __restrict__ st, const int* __restrict__ d_count,
                                                unsigned long long* __restrict__ tot)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long a = 0, b = 0, c = 0;
    if (i < *d_count) {
        a = (unsigned)st[3 * i];
        b = (unsigned)st[3 * i + 1];
        c = (unsigned)st[3 * i + 2];
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        c += __shfl_xor(c, o);
    }
    if (__lane_id() == 0 && (a | b | c)) {
        atomicAdd(&tot[0], a);
        atomicAdd(&tot[1], b);
        atomicAdd(&tot[2], c);
    }
}

} // namespace

extern "C" int32_t mt_render(mt_context* ctx, const mt_scene* scene, const float* materials,
                             const mt_envmap* env, const int32_t* light_faces, const int32_t* d_num_lights,
                             const mt_render_desc* desc, float* image, mt_render_stats* stats)
{
    MT_ARG_CHECK(ctx && scene && materials && desc && image, "mt_render: NULL argument");
    MT_ARG_CHECK(desc->width > 0 && desc->height > 0 && desc->spp > 0, "mt_render: empty image");
    MT_ARG_CHECK(desc->max_bounces >= 1 && desc->max_bounces <= 63, "mt_render: max_bounces %d not in [1,63]",
                 desc->max_bounces);
    MT_ARG_CHECK(!light_faces == !d_num_lights, "mt_render: light_faces and d_num_lights go together");
    int64_t npix = (int64_t)desc->width * desc->height;
    int64_t np64 = npix * desc->spp;
    MT_ARG_CHECK(np64 <= (1ll << 30), "mt_render: %lld paths exceed the 2^30 wavefront limit", (long long)np64);
    int np = (int)np64;
    hipStream_t st = ctx->stream;

    // ---- camera basis (host, double)
    Camera C;
    {
        double f[3] = {(double)desc->look[0] - desc->pos[0], (double)desc->look[1] - desc->pos[1],
                       (double)desc->look[2] - desc->pos[2]};
        double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        MT_ARG_CHECK(fl > 0.0, "mt_render: camera look-at equals position");
        for (double& x : f) x /= fl;
        const float* up = desc->up;
        double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
        double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        MT_ARG_CHECK(rl > 0.0, "mt_render: up vector parallel to the view direction");
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
    }

    // ---- buffers (one allocation)
    const int nb = desc->max_bounces;
    size_t n = (size_t)np;
    size_t bytes = 0;
    auto take = [&](size_t b) { size_t off = bytes; bytes += (b + 255) & ~(size_t)255; return off; };
    size_t oT = take(16 * n), oL = take(16 * n);
    size_t oA_o = take(16 * n), oA_d = take(16 * n), oB_o = take(16 * n), oB_d = take(16 * n);
    size_t oQA = take(4 * n), oQB = take(4 * n);
    size_t oPrim = take(4 * n), oT_ = take(4 * n), oAlive = take(4 * n), oOcc = take(4 * n);
    size_t oSo = take(16 * n), oSd = take(16 * n), oSc = take(16 * n);
    size_t oCnt = take(4 * 64), oShc = take(4 * 64);
    const bool want_ilp = (desc->flags & MT_RENDER_TRAVERSAL_STATS) != 0;
    const bool profile = (desc->flags & MT_RENDER_PROFILE) != 0;
    size_t oIlp = want_ilp ? take(12 * n) : 0, oTot = take(8 * 8);
    char* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) {
        mt_set_error("mt_render: cannot allocate %.2f GB of wavefront state", bytes / 1e9);
        return MT_ERR_ALLOC;
    }
    auto F4 = [&](size_t off) { return reinterpret_cast<float4*>(base + off); };
    auto I = [&](size_t off) { return reinterpret_cast<int*>(base + off); };
    PathState S{F4(oT), F4(oL)};
    float4* ro[2] = {F4(oA_o), F4(oB_o)};
    float4* rd[2] = {F4(oA_d), F4(oB_d)};
    int* q[2] = {I(oQA), I(oQB)};
    int* counts = I(oCnt);
    int* shc = I(oShc);
    int32_t rc = MT_OK;
    auto chk = [&](void) -> int32_t {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            mt_set_error("mt_render: launch failed: %s", hipGetErrorString(e));
            return MT_ERR_HIP;
        }
        return MT_OK;
    };

    SceneView V;
    V.points = reinterpret_cast<const float4*>(scene->desc.points);
    V.normals = reinterpret_cast<const float4*>(scene->desc.normals);
    V.faces = scene->desc.faces;
    V.mats = reinterpret_cast<const MtMaterial*>(materials);
    V.light_faces = light_faces;
    V.num_lights = d_num_lights;
    V.env = env_view(env);

    int grid = mt_blocks(np, RT);
    const int gshade = grid, gnee = grid; // one slot per lane; empty work-groups retire at once
    unsigned long long* tot = reinterpret_cast<unsigned long long*>(base + oTot);
    int* ilp = want_ilp ? I(oIlp) : nullptr;
    hipEvent_t ev[4 * 64];
    int nev = 0;
    if (profile) {
        for (; nev < 4 * nb; nev++)
            if (hipEventCreate(&ev[nev]) != hipSuccess) {
                rc = MT_ERR_HIP;
                mt_set_error("mt_render: hipEventCreate failed");
                break;
            }
    }
    k_zero_i32<<<1, 128, 0, st>>>(counts, 128); // counts[64] and shadow counts[64] are adjacent
    k_zero_i32<<<1, 64, 0, st>>>(reinterpret_cast<int*>(tot), 16);
    k_camera<<<grid, RT, 0, st>>>(C, (int)npix, np, desc->seed, ro[0], rd[0], q[0], S, counts);
    if (rc == MT_OK) rc = chk();
    int cur = 0;
    for (int b = 0; b < nb && rc == MT_OK; b++) {
        int* dcount = counts + b;
        if (profile) hipEventRecord(ev[4 * b], st);
        rc = mt_trace_closest(ctx, scene, reinterpret_cast<float*>(ro[cur]), reinterpret_cast<float*>(rd[cur]),
                              dcount, np, I(oPrim), reinterpret_cast<float*>(base + oT_), ilp, desc->mode);
        if (rc) break;
        if (profile) hipEventRecord(ev[4 * b + 1], st);
        if (want_ilp) k_sum_ilp<<<grid, RT, 0, st>>>(ilp, dcount, tot);
        ShadeArgs A;
        A.V = V;
        A.S = S;
        A.queue = q[cur];
        A.d_count = dcount;
        A.o_tmin = ro[cur];
        A.d_tmax = rd[cur];
        A.prim = I(oPrim);
        A.tin = reinterpret_cast<const float*>(base + oT_);
        A.alive = I(oAlive);
        A.s_o_tmin = F4(oSo);
        A.s_d_tmax = F4(oSd);
        A.s_contrib = F4(oSc);
        A.shadow_count = shc + b;
        A.bounce = b;
        A.rr_start = desc->rr_start;
        A.flags = desc->flags;
        A.seed = desc->seed;
        k_shade<<<gshade, RT, 0, st>>>(A);
        if ((rc = chk())) break;
        if (profile) hipEventRecord(ev[4 * b + 2], st);
        rc = mt_trace_any(ctx, scene, reinterpret_cast<float*>(F4(oSo)), reinterpret_cast<float*>(F4(oSd)), dcount,
                          np, I(oOcc), ilp, desc->mode);
        if (rc) break;
        if (profile) hipEventRecord(ev[4 * b + 3], st);
        if (want_ilp) k_sum_ilp<<<grid, RT, 0, st>>>(ilp, dcount, tot + 3);
        k_nee_accum<<<gnee, RT, 0, st>>>(q[cur], dcount, F4(oSc), I(oOcc), S);
        if ((rc = chk())) break;
        if (b + 1 < nb) {
            mt_payload P[3] = {{q[cur], q[1 - cur], 4}, {ro[cur], ro[1 - cur], 16}, {rd[cur], rd[1 - cur], 16}};
            rc = mt_compact(ctx, I(oAlive), 0, dcount, np, nullptr, counts + b + 1, P, 3);
            if (rc) break;
            cur = 1 - cur;
        }
    }
    if (rc == MT_OK) {
        k_resolve<<<mt_blocks(npix, RT), RT, 0, st>>>(S.L, (int)npix, desc->spp, reinterpret_cast<float4*>(image));
        rc = chk();
    }
    if (rc == MT_OK && stats) {
        int hc[128];
        unsigned long long ht[6];
        hipError_t e = hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(ht, tot, sizeof(ht), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            mt_set_error("mt_render: stats readback failed: %s", hipGetErrorString(e));
            rc = MT_ERR_HIP;
        } else {
            stats->extension_rays = 0;
            stats->shadow_rays = 0;
            stats->paths = np;
            for (int b = 0; b < 64; b++) stats->bounce_counts[b] = b < nb ? hc[b] : 0;
            for (int b = 0; b < nb; b++) {
                stats->extension_rays += hc[b];
                stats->shadow_rays += hc[64 + b];
            }
            for (int k = 0; k < 3; k++) {
                stats->closest_ilp[k] = (int64_t)ht[k];
                stats->any_ilp[k] = (int64_t)ht[3 + k];
            }
            stats->closest_launches = stats->any_launches = 0;
            stats->closest_ms = stats->any_ms = 0.0f;
            if (profile) {
                for (int b = 0; b < nb; b++) {
                    float ms = 0.0f;
                    if (hipEventElapsedTime(&ms, ev[4 * b], ev[4 * b + 1]) == hipSuccess) {
                        stats->closest_ms += ms;
                        stats->closest_launches++;
                    }
                    if (hipEventElapsedTime(&ms, ev[4 * b + 2], ev[4 * b + 3]) == hipSuccess) {
                        stats->any_ms += ms;
                        stats->any_launches++;
                    }
                }
            }
        }
    }
    hipStreamSynchronize(st);
    for (int k = 0; k < nev; k++) hipEventDestroy(ev[k]);
    hipFree(base);
    if (rc == MT_OK) rc = mt_scan_check(ctx);
    return rc;
}
