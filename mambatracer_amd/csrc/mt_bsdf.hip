// Stage B batch entry points: setupBsdf / SampleBsdf / EvaluateBsdf over n
// records (src/cl/kernel/Bsdf.cl:130-160,525-606).  The wavefront render
// loop calls the same device functions (mt_bsdf.h) fused into its kernels;
// these entry points expose them one call per stage, as the reference's
// SetupBSDFPath / SampleBSDFRayDirection / DirectLight kernels do
// (Render.cl:10-40,233-285,377-466).
//
// BSDF record (MT_BSDF_WORDS = 20 floats): frame mX, mY, mZ (9), localDirFix
// (3), localGeomN (3), diffProb, glossyProb, 3 reserved; material id apart.
#include "mt_bsdf.h"
#include "mt_internal.h"

namespace {

__device__ __forceinline__ MtBsdf load_bsdf(const float* __restrict__ rec, int mat)
{
    MtBsdf B;
    B.F.mx = mk3(rec[0], rec[1], rec[2]);
    B.F.my = mk3(rec[3], rec[4], rec[5]);
    B.F.mz = mk3(rec[6], rec[7], rec[8]);
    B.fix = mk3(rec[9], rec[10], rec[11]);
    B.geom = mk3(rec[12], rec[13], rec[14]);
    B.diff_prob = rec[15];
    B.glossy_prob = rec[16];
    B.mat = mat;
    return B;
}

__global__ void k_bsdf_setup(int n, const float4* __restrict__ ray_d, const float4* __restrict__ isn,
                             const float4* __restrict__ isng, const int* __restrict__ mat,
                             const MtMaterial* __restrict__ mats, float* __restrict__ out,
                             int* __restrict__ mat_id)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    MtBsdf B = mt_bsdf_setup_one(xyz(ray_d[i]), xyz(isn[i]), xyz(isng[i]), mat[i], mats);
    float* r = out + (size_t)MT_BSDF_WORDS * i;
    r[0] = B.F.mx.x; r[1] = B.F.mx.y; r[2] = B.F.mx.z;
    r[3] = B.F.my.x; r[4] = B.F.my.y; r[5] = B.F.my.z;
    r[6] = B.F.mz.x; r[7] = B.F.mz.y; r[8] = B.F.mz.z;
    r[9] = B.fix.x; r[10] = B.fix.y; r[11] = B.fix.z;
    r[12] = B.geom.x; r[13] = B.geom.y; r[14] = B.geom.z;
    r[15] = B.diff_prob; r[16] = B.glossy_prob;
    r[17] = 0.0f; r[18] = 0.0f; r[19] = 0.0f;
    mat_id[i] = B.mat;
}

__global__ void k_bsdf_sample(int n, const float* __restrict__ rec, const int* __restrict__ mat_id,
                              const MtMaterial* __restrict__ mats, const float* __restrict__ u,
                              float* __restrict__ f, float* __restrict__ dir,
                              float* __restrict__ pdf, float* __restrict__ cos_out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int m = mat_id[i];
    MtBsdf B = load_bsdf(rec + (size_t)MT_BSDF_WORDS * i, m);
    const MtMaterial& M = mats[m >= 0 ? m : 0];
    v3 wd;
    float p = 0.0f, c = 0.0f;
    v3 fv = mt_bsdf_sample_one(B, M, u[3 * i], u[3 * i + 1], u[3 * i + 2], wd, p, c);
    f[3 * i] = fv.x; f[3 * i + 1] = fv.y; f[3 * i + 2] = fv.z;
    dir[3 * i] = wd.x; dir[3 * i + 1] = wd.y; dir[3 * i + 2] = wd.z;
    pdf[i] = p;
    cos_out[i] = c;
}

__global__ void k_bsdf_eval(int n, const float* __restrict__ rec, const int* __restrict__ mat_id,
                            const MtMaterial* __restrict__ mats, const float* __restrict__ dir,
                            float* __restrict__ f, float* __restrict__ pdf,
                            float* __restrict__ cos_out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int m = mat_id[i];
    MtBsdf B = load_bsdf(rec + (size_t)MT_BSDF_WORDS * i, m);
    const MtMaterial& M = mats[m >= 0 ? m : 0];
    float p = 0.0f, c = 0.0f;
    v3 fv = mt_bsdf_eval_one(B, M, mk3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]), p, c);
    f[3 * i] = fv.x; f[3 * i + 1] = fv.y; f[3 * i + 2] = fv.z;
    pdf[i] = p;
    cos_out[i] = c;
}

} // namespace

extern "C" int32_t mt_bsdf_setup(mt_context* ctx, int32_t n, const float* ray_d, const float* isect_n,
                                 const float* isect_ng, const int32_t* mat, const float* materials,
                                 float* bsdf, int32_t* mat_id)
{
    MT_ARG_CHECK(ctx && ray_d && isect_n && isect_ng && mat && materials && bsdf && mat_id,
                 "mt_bsdf_setup: NULL argument");
    MT_ARG_CHECK(n >= 0, "mt_bsdf_setup: negative n");
    if (n == 0) return MT_OK;
    k_bsdf_setup<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(
        n, reinterpret_cast<const float4*>(ray_d), reinterpret_cast<const float4*>(isect_n),
        reinterpret_cast<const float4*>(isect_ng), mat, reinterpret_cast<const MtMaterial*>(materials),
        bsdf, mat_id);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_bsdf_sample(mt_context* ctx, int32_t n, const float* bsdf, const int32_t* mat_id,
                                  const float* materials, const float* u, float* f, float* dir,
                                  float* pdf, float* cos_out)
{
    MT_ARG_CHECK(ctx && bsdf && mat_id && materials && u && f && dir && pdf && cos_out,
                 "mt_bsdf_sample: NULL argument");
    MT_ARG_CHECK(n >= 0, "mt_bsdf_sample: negative n");
    if (n == 0) return MT_OK;
    k_bsdf_sample<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(
        n, bsdf, mat_id, reinterpret_cast<const MtMaterial*>(materials), u, f, dir, pdf, cos_out);
    MT_LAUNCH_CHECK();
    return MT_OK;
}

extern "C" int32_t mt_bsdf_eval(mt_context* ctx, int32_t n, const float* bsdf, const int32_t* mat_id,
                                const float* materials, const float* dir, float* f, float* pdf,
                                float* cos_out)
{
    MT_ARG_CHECK(ctx && bsdf && mat_id && materials && dir && f && pdf && cos_out,
                 "mt_bsdf_eval: NULL argument");
    MT_ARG_CHECK(n >= 0, "mt_bsdf_eval: negative n");
    if (n == 0) return MT_OK;
    k_bsdf_eval<<<mt_blocks(n, 256), 256, 0, ctx->stream>>>(
        n, bsdf, mat_id, reinterpret_cast<const MtMaterial*>(materials), dir, f, pdf, cos_out);
    MT_LAUNCH_CHECK();
    return MT_OK;
}
