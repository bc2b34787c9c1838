// Stage B device functions: GGX (anisotropic Trowbridge-Reitz) + Lambert
// BSDF setup, sampling and evaluation, and the reference RNG.
//
// Restates src/cl/kernel/Bsdf.cl:70-606, Geometry.cl:1-83,
// Sampling.cl:15-59,129-148 and Util.cl:194-221 in the evaluation order the
// file header of mt_devmath.h pins.  Reference quirks reproduced:
//  - the glossy pdf is pdfGGXVNDF(wo, wh): the generated direction is used as
//    the view vector (Bsdf.cl:417,492);
//  - M_1_PI expands unparenthesised (Util.cl:8): x * M_1_PI == (x * 1) / pi;
//  - sampleGGXVNDF's double literals (Bsdf.cl:282,285-286) are evaluated in
//    double;
//  - the isFloat4Zero test in SampleBsdf never fires because .w != 0
//    (Bsdf.cl:580,589).
// Deviations: rsqrt (Bsdf.cl:277) is 1/sqrtf; the sampled pdf starts at 0
// (SampleBSDFRayDirection passes it uninitialised, Render.cl:262-263).
#pragma once

#include "mt_devmath.h"

struct MtFrame {
    v3 mx, my, mz;
};

struct MtBsdf {
    MtFrame F;
    v3 fix;   // localDirFix: local(-ray.d)
    v3 geom;  // localGeomN
    float diff_prob, glossy_prob;
    int mat;  // -1: invalid (grazing incidence)
};

// material record: 4 int flags + 8 float4 (Bsdf.cl:24-42), 144 bytes
struct MtMaterial {
    int flags[4];
    float4 diffuse_color, diffuse_param, glossy_color, glossy_param;
    float4 mirror_color, mirror_param, emission_color, emission_param;
};

__device__ __forceinline__ MtFrame mt_get_frame(v3 n)
{ // Util.cl:196-211
    MtFrame F;
    F.mz = n;
    float a = 1.0f / (1.0f + n.z);
    float b = (-n.x * n.y) * a;
    F.mx = mk3(1.0f - (n.x * n.x) * a, b, -n.x);
    F.my = mk3(b, 1.0f - (n.y * n.y) * a, -n.y);
    if (n.z < -0.9999999f) {
        F.mx = mk3(0.0f, -1.0f, 0.0f);
        F.my = mk3(-1.0f, 0.0f, 0.0f);
    }
    return F;
}

__device__ __forceinline__ v3 mt_to_local(const MtFrame& F, v3 a)
{ // local_coordinate, Util.cl:218-221
    return mk3(dot3(F.mx, a), dot3(F.my, a), dot3(F.mz, a));
}

__device__ __forceinline__ v3 mt_to_world(const MtFrame& F, v3 a)
{ // world_coordinate, Util.cl:213-216
    return add3(add3(scl3(F.mx, a.x), scl3(F.my, a.y)), scl3(F.mz, a.z));
}

// Geometry.cl:236-291
__device__ __forceinline__ float g_cos2theta(v3 w) { return w.z * w.z; }
__device__ __forceinline__ float g_sin2theta(v3 w) { return fmaxf(0.0f, 1.0f - g_cos2theta(w)); }
__device__ __forceinline__ float g_sintheta(v3 w) { return sqrtf(g_sin2theta(w)); }
__device__ __forceinline__ float g_tantheta(v3 w) { return g_sintheta(w) / w.z; }
__device__ __forceinline__ float g_tan2theta(v3 w) { return g_sin2theta(w) / g_cos2theta(w); }
__device__ __forceinline__ float g_clamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float g_cosphi(v3 w)
{
    float s = g_sintheta(w);
    return s == 0.0f ? 1.0f : g_clamp(w.x / s, -1.0f, 1.0f);
}
__device__ __forceinline__ float g_sinphi(v3 w)
{
    float s = g_sintheta(w);
    return s == 0.0f ? 0.0f : g_clamp(w.y / s, -1.0f, 1.0f);
}
__device__ __forceinline__ float g_cos2phi(v3 w) { float c = g_cosphi(w); return c * c; }
__device__ __forceinline__ float g_sin2phi(v3 w) { float s = g_sinphi(w); return s * s; }

__device__ __forceinline__ float ggx_lambda(v3 w, float ax, float ay)
{ // alpha + lambda, Bsdf.cl:229-242
    float abs_tan = fabsf(g_tantheta(w));
    float a = sqrtf((g_cos2phi(w) * ax) * ax + (g_sin2phi(w) * ay) * ay);
    float a2t2 = (a * abs_tan) * (a * abs_tan);
    float lam = (-1.0f + sqrtf(1.0f + a2t2)) / 2.0f;
    return isinf(abs_tan) ? 0.0f : lam;
}

__device__ __forceinline__ float ggx_D(v3 wh, float ax, float ay)
{ // D_H, Bsdf.cl:245-253
    float t2 = g_tan2theta(wh);
    float c4 = g_cos2theta(wh) * g_cos2theta(wh);
    float e = (g_cos2phi(wh) / (ax * ax) + g_sin2phi(wh) / (ay * ay)) * t2;
    float d = 1.0f / (((((MT_PI * ax) * ay) * c4) * (1.0f + e)) * (1.0f + e));
    return isinf(t2) ? 0.0f : d;
}

__device__ __forceinline__ float ggx_G(v3 w1, v3 w2, float ax, float ay)
{ // G_Atten, Bsdf.cl:256-259
    return 1.0f / ((1.0f + ggx_lambda(w1, ax, ay)) + ggx_lambda(w2, ax, ay));
}

__device__ __forceinline__ float ggx_G1(v3 w, float ax, float ay)
{
    return 1.0f / (1.0f + ggx_lambda(w, ax, ay));
}

__device__ __forceinline__ v3 ggx_sample_vndf(v3 v, float ax, float ay, float r1, float r2)
{ // sampleGGXVNDF, Bsdf.cl:270-294
    v3 vh = norm3(mk3(ax * v.x, ay * v.y, v.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    v3 t1v;
    if (lensq > 0.0f) {
        float inv = 1.0f / sqrtf(lensq);
        t1v = mk3(-vh.y * inv, vh.x * inv, 0.0f * inv);
    } else {
        t1v = mk3(1.0f, 0.0f, 0.0f);
    }
    v3 t2v = cross3(vh, t1v);
    float r = sqrtf(r1);
    float phi = (float)((2.0 * (double)MT_PI) * (double)r2);
    float t1 = r * mt_xm_cosf(phi);
    float t2 = r * mt_xm_sinf(phi);
    float s = (float)(0.5 * (1.0 + (double)vh.z));
    t2 = (float)((1.0 - (double)s) * sqrt(1.0 - (double)(t1 * t1)) + (double)s * (double)t2);
    float k = sqrtf(fmaxf(0.0f, (1.0f - t1 * t1) - t2 * t2));
    v3 nh = add3(add3(scl3(t1v, t1), scl3(t2v, t2)), scl3(vh, k));
    return norm3(mk3(ax * nh.x, ay * nh.y, fmaxf(0.0f, nh.z)));
}

__device__ __forceinline__ float ggx_pdf(v3 v, v3 ni, float ax, float ay)
{ // VNDF + pdfGGXVNDF, Bsdf.cl:297-311
    float dv = ((ggx_G1(v, ax, ay) * fmaxf(0.001f, dot3(v, ni))) * ggx_D(ni, ax, ay)) / fmaxf(0.001f, v.z);
    float vn = dot3(v, ni);
    return dv / (4.0f * vn);
}

__device__ __forceinline__ MtBsdf mt_bsdf_setup_one(v3 ray_d, v3 n, v3 ng, int mat,
                                                     const MtMaterial* __restrict__ mats)
{ // setupBsdf, Bsdf.cl:130-160
    MtBsdf B;
    B.F = mt_get_frame(n);
    B.fix = mt_to_local(B.F, neg3(ray_d));
    B.geom = mk3(0, 0, 0);
    B.diff_prob = 0.0f;
    B.glossy_prob = 0.0f;
    if (fabsf(B.fix.z) < MT_EPS_COSINE || mat < 0) { // grazing, or no surface (a miss)
        B.mat = -1;
        return B;
    }
    B.geom = mt_to_local(B.F, ng);
    B.mat = mat;
    const MtMaterial& m = mats[mat];
    // GetComponentProbabilities, Bsdf.cl:93-128 (AlbedoDiffuse/Glossy :70-80)
    float s1 = m.diffuse_param.x, s2 = m.glossy_param.x;
    float ad = luminance3(m.diffuse_color.x * s1, m.diffuse_color.y * s1, m.diffuse_color.z * s1);
    float ag = luminance3(m.glossy_color.x * s2, m.glossy_color.y * s2, m.glossy_color.z * s2);
    float tot = ad + ag;
    if (!(tot < 1e-9f)) {
        B.diff_prob = ad / tot;
        B.glossy_prob = ag / tot;
    }
    return B;
}

__device__ __forceinline__ v3 mt_schlick(float c, float4 f0)
{ // SchlickColor, Bsdf.cl:185-189
    float e = mt_xm_pow5f(1.0f - c);
    return mk3(f0.x + (1.0f - f0.x) * e, f0.y + (1.0f - f0.y) * e, f0.z + (1.0f - f0.z) * e);
}

// Lambert value c * M_1_PI == (c * 1) / pi (Bsdf.cl:393,466)
__device__ __forceinline__ v3 mt_lambert(const MtMaterial& m)
{
    return mk3((m.diffuse_color.x * 1.0f) / MT_PI, (m.diffuse_color.y * 1.0f) / MT_PI,
               (m.diffuse_color.z * 1.0f) / MT_PI);
}

__device__ __forceinline__ v3 mt_glossy_value(const MtMaterial& m, v3 wi, v3 wo, v3 wh, float ax,
                                              float ay)
{ // shared tail of SampleGlossy/EvaluateGlossy, Bsdf.cl:414-430,489-505
    float ci = fabsf(wi.z), co = fabsf(wo.z);
    float fr = (ggx_G(wi, wo, ax, ay) * ggx_D(wh, ax, ay)) / ((4.0f * ci) * co);
    float whdot = fabsf(dot3(wh, wo));
    v3 fc = mt_schlick(whdot, m.glossy_color);
    if (ci == 0.0f || co == 0.0f) fc = mk3(0, 0, 0);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) fc = mk3(0, 0, 0);
    return mk3(fc.x * fr, fc.y * fr, fc.z * fr);
}

// EvaluateDiffuse, Bsdf.cl:452-467; returns whether the lobe contributed (w = 1)
__device__ __forceinline__ bool mt_eval_diffuse(const MtBsdf& B, const MtMaterial& m, v3 gen,
                                                float& pdf, v3& f)
{
    if (B.diff_prob == 0.0f) return false;
    if (B.fix.z < MT_EPS_COSINE || gen.z < MT_EPS_COSINE) return false;
    pdf += B.diff_prob * fmaxf(0.0f, (gen.z * 1.0f) / MT_PI);
    f = add3(f, mt_lambert(m));
    return true;
}

// EvaluateGlossy, Bsdf.cl:469-507
__device__ __forceinline__ void mt_eval_glossy(const MtBsdf& B, const MtMaterial& m, v3 gen,
                                               float& pdf, v3& f)
{
    if (B.glossy_prob == 0.0f) return;
    if (B.fix.z < MT_EPS_COSINE || gen.z < MT_EPS_COSINE) return;
    float ax = m.glossy_param.y, ay = m.glossy_param.z;
    v3 wh = norm3(add3(B.fix, gen)); // halfVector, Bsdf.cl:336-341
    pdf += B.glossy_prob * ggx_pdf(gen, wh, ax, ay);
    f = add3(f, mt_glossy_value(m, B.fix, gen, wh, ax, ay));
}

// EvaluateBsdf, Bsdf.cl:525-545.  pdf and cos are accumulated/overwritten as
// in the reference (DirectLight starts both at 0, Render.cl:440).
__device__ __forceinline__ v3 mt_bsdf_eval_one(const MtBsdf& B, const MtMaterial& m, v3 wdir,
                                               float& pdf, float& cos_out)
{
    v3 f = mk3(0, 0, 0);
    v3 gen = mt_to_local(B.F, wdir);
    if (gen.z * B.fix.z < 0.0f) return f;
    cos_out = gen.z;
    mt_eval_diffuse(B, m, gen, pdf, f);
    mt_eval_glossy(B, m, gen, pdf, f);
    return f;
}

// SampleBsdf, Bsdf.cl:555-606 (with SampleDiffuse :380-394, SampleGlossy
// :396-432, SampleCosHemisphereW Sampling.cl:129-148).  Returns f (zero when
// the sample is rejected); wdir/pdf/cos as the reference leaves them.
// Either pick evaluates the glossy lobe at a (direction, half vector) pair:
// the diffuse pick at its cosine-sampled direction and their half vector
// (EvaluateGlossy), the glossy pick at the VNDF half vector and its
// reflection.  Lanes of one wave draw both picks (u2 per path), so the pair
// is formed per pick and the lobe's pdf and value computed once after it,
// on the same values either pick passes; the sums then follow each pick's
// order (diffuse pick: diffuse term, then glossy; glossy pick: glossy, then
// EvaluateDiffuse's term).
__device__ __forceinline__ v3 mt_bsdf_sample_one(const MtBsdf& B, const MtMaterial& m, float u0,
                                                 float u1, float u2, v3& wdir, float& pdf,
                                                 float& cos_out)
{
    const bool dsel = u2 < B.diff_prob;
    const bool front = B.fix.z >= MT_EPS_COSINE;
    const float ax = m.glossy_param.y, ay = m.glossy_param.z;
    v3 gen = mk3(0, 0, 0), wh = mk3(0, 0, 0);
    if (front) {
        if (dsel) {
            float term1 = (2.0f * MT_PI) * u0;
            float term2 = sqrtf(1.0f - u1);
            gen = mk3(mt_xm_cosf(term1) * term2, mt_xm_sinf(term1) * term2, sqrtf(u1));
        } else {
            wh = ggx_sample_vndf(B.fix, ax, ay, u0, u1);
            float k = 2.0f * dot3(B.fix, wh);
            gen = add3(neg3(B.fix), scl3(wh, k)); // reflectFromN, Bsdf.cl:325-328
        }
    }
    // the evaluation functions' own test (mt_eval_diffuse / mt_eval_glossy)
    const bool eval_ok = !(B.fix.z < MT_EPS_COSINE || gen.z < MT_EPS_COSINE);
    const bool g_on = dsel ? (B.glossy_prob != 0.0f && eval_ok) : front;
    float gpdf = 0.0f;
    v3 gval = mk3(0, 0, 0);
    if (g_on) {
        if (dsel) wh = norm3(add3(B.fix, gen)); // halfVector, Bsdf.cl:336-341
        gpdf = B.glossy_prob * ggx_pdf(gen, wh, ax, ay);
        gval = mt_glossy_value(m, B.fix, gen, wh, ax, ay);
    }
    v3 f = mk3(0, 0, 0);
    if (dsel) {
        if (front) {
            pdf += ((gen.z * 1.0f) / MT_PI) * B.diff_prob;
            f = mt_lambert(m);
        }
        if (g_on) {
            pdf += gpdf;
            f = add3(f, gval);
        }
    } else {
        if (front) {
            pdf += gpdf;
            f = gval;
        }
        if (B.diff_prob != 0.0f && eval_ok) {
            pdf += B.diff_prob * fmaxf(0.0f, (gen.z * 1.0f) / MT_PI);
            f = add3(f, mt_lambert(m));
        }
    }
    float c = fabsf(gen.z);
    cos_out = c;
    if (c < MT_EPS_COSINE) {
        wdir = mk3(0, 0, 0);
        return mk3(0, 0, 0);
    }
    wdir = mt_to_world(B.F, gen);
    return f;
}

// ---- RNG, Sampling.cl:15-59 (32-bit wrapping, arithmetic shifts) ----
__device__ __forceinline__ int32_t mt_wang_hash(int32_t seed)
{
    int32_t s = (seed ^ 61) ^ (seed >> 16);
    s = (int32_t)((uint32_t)s * 9u);
    s = s ^ (s >> 4);
    s = (int32_t)((uint32_t)s * 0x27d4eb2du);
    s = s ^ (s >> 15);
    return s;
}

__device__ __forceinline__ float mt_get_random(int32_t& x, int32_t& y)
{
    x = mt_wang_hash(x);
    y = mt_wang_hash(y);
    uint32_t ires = ((uint32_t)x << 16) + (uint32_t)y;
    return (__uint_as_float((ires & 0x007fffffu) | 0x40000000u) - 2.0f) / 2.0f;
}
