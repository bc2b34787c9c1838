/*
 * ORACLE (test infrastructure only) — stage B: GGX/Lambert BSDF setup,
 * sampling and evaluation, and the reference RNG.
 *
 * Parity status
 *   orc_bsdf_setup   restates setupBsdf + GetComponentProbabilities + get_frame
 *                    (Bsdf.cl:93-160, Util.cl:194-221).  Pinned by frame
 *                    orthonormality and local(-d).z == dot(n, -d).
 *   orc_ggx_*        restate D_H, lambda, sampleGGXVNDF, pdfGGXVNDF
 *                    (Bsdf.cl:229-311, Geometry.cl:1-83).  Pinned against the
 *                    textbook forms: normalisation integral of D(h)cos(h) == 1,
 *                    G1 <= 1, isotropic Smith lambda closed form, and a
 *                    Monte-Carlo histogram of sampleGGXVNDF against
 *                    D_v(h) = G1(v) max(0, v.h) D(h) / v.z (Heitz 2018).
 *   orc_bsdf_sample  restates SampleBsdf/SampleDiffuse/SampleGlossy
 *                    (Bsdf.cl:380-432,555-606, Sampling.cl:129-148).
 *   orc_bsdf_eval    restates EvaluateBsdf/EvaluateDiffuse/EvaluateGlossy
 *                    (Bsdf.cl:452-545).  Pinned by reciprocity of f
 *                    (f(wi,wo) == f(wo,wi)), the Lambert closed form c/pi and
 *                    the sample/eval consistency of f.
 *   The glossy pdf is evaluated with the generated direction as view vector,
 *   pdfGGXVNDF(wo, wh) (Bsdf.cl:417,492); this reference quirk is reproduced
 *   and is pinned only by the restatement (parity of the quirk itself is
 *   "restatement-pinned": no textbook form computes it).
 *   orc_wang_hash / orc_get_random restate Sampling.cl:15-59 with 32-bit
 *   wrapping arithmetic and arithmetic right shifts; pinned by hand-computed
 *   known answers in tests/test_oracle_bsdf.py.
 *
 * Deviations (documented): SampleBSDFRayDirection passes an uninitialised pdf
 * into SampleBsdf, which only adds to it (Render.cl:262-263, Bsdf.cl:391,417);
 * the oracle starts it at 0.  rsqrt (Bsdf.cl:277) is pinned as 1/sqrtf.
 * Double-precision literals in sampleGGXVNDF (Bsdf.cl:282,285-286) are
 * evaluated in double, as OpenCL does on fp64-capable devices.
 */
#include "orc.h"
#include "orc_math.h"

#define EPS_COSINE 1e-6f

typedef struct {
    f4 mx, my, mz;
} oframe;

static oframe get_frame(f4 n)
{ /* Util.cl:196-211 */
    oframe F;
    F.mz = n;
    float a = 1.0f / (1.0f + n.z);
    float b = (-n.x * n.y) * a;
    F.mx = f4_make(1.0f - (n.x * n.x) * a, b, -n.x, 0.0f);
    F.my = f4_make(b, 1.0f - (n.y * n.y) * a, -n.y, 0.0f);
    if (n.z < -0.9999999f) {
        F.mx = f4_make(0.0f, -1.0f, 0.0f, 0.0f);
        F.my = f4_make(-1.0f, 0.0f, 0.0f, 0.0f);
    }
    return F;
}

static f4 to_local(const oframe* F, f4 a)
{ /* local_coordinate, Util.cl:218-221 */
    return f4_make(dot3(F->mx, a), dot3(F->my, a), dot3(F->mz, a), 0.0f);
}

static f4 to_world(const oframe* F, f4 a)
{ /* world_coordinate, Util.cl:213-216 */
    return f4_add(f4_add(f4_scale(F->mx, a.x), f4_scale(F->my, a.y)), f4_scale(F->mz, a.z));
}

/* Geometry.cl:236-291 (file lines 1-56) */
static float cos2theta(f4 w) { return w.z * w.z; }
static float sin2theta(f4 w) { return fmaxf(0.0f, 1.0f - cos2theta(w)); }
static float sintheta(f4 w) { return sqrtf(sin2theta(w)); }
static float tantheta(f4 w) { return sintheta(w) / w.z; }
static float tan2theta(f4 w) { return sin2theta(w) / cos2theta(w); }
static float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static float cosphi(f4 w) { float s = sintheta(w); return s == 0.0f ? 1.0f : clampf(w.x / s, -1.0f, 1.0f); }
static float sinphi(f4 w) { float s = sintheta(w); return s == 0.0f ? 0.0f : clampf(w.y / s, -1.0f, 1.0f); }
static float cos2phi(f4 w) { return cosphi(w) * cosphi(w); }
static float sin2phi(f4 w) { return sinphi(w) * sinphi(w); }

static float ggx_alpha(f4 w, float ax, float ay)
{ /* Bsdf.cl:229-232 */
    return sqrtf((cos2phi(w) * ax) * ax + (sin2phi(w) * ay) * ay);
}

static float ggx_lambda(f4 w, float ax, float ay)
{ /* Bsdf.cl:235-242 */
    float abs_tan = fabsf(tantheta(w));
    float a = ggx_alpha(w, ax, ay);
    float a2t2 = (a * abs_tan) * (a * abs_tan);
    float lam = (-1.0f + sqrtf(1.0f + a2t2)) / 2.0f;
    return isinf(abs_tan) ? 0.0f : lam;
}

static float ggx_D(f4 wh, float ax, float ay)
{ /* Bsdf.cl:245-253 */
    float t2 = tan2theta(wh);
    float c4 = cos2theta(wh) * cos2theta(wh);
    float e = (cos2phi(wh) / (ax * ax) + sin2phi(wh) / (ay * ay)) * t2;
    float d = 1.0f / (((((ORC_PI * ax) * ay) * c4) * (1.0f + e)) * (1.0f + e));
    return isinf(t2) ? 0.0f : d;
}

static float ggx_G(f4 w1, f4 w2, float ax, float ay)
{ /* G_Atten, Bsdf.cl:256-259 */
    return 1.0f / ((1.0f + ggx_lambda(w1, ax, ay)) + ggx_lambda(w2, ax, ay));
}

static float ggx_G1(f4 w, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(w, ax, ay)); }

static f4 ggx_sample_vndf(f4 v, float ax, float ay, float r1, float r2)
{ /* sampleGGXVNDF, Bsdf.cl:270-294 */
    f4 vh = normalize3(f4_make(ax * v.x, ay * v.y, v.z, 0.0f));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    f4 t1v;
    if (lensq > 0.0f) {
        float inv = 1.0f / sqrtf(lensq);
        t1v = f4_make(-vh.y * inv, vh.x * inv, 0.0f * inv, 0.0f);
    } else {
        t1v = f4_make(1.0f, 0.0f, 0.0f, 0.0f);
    }
    f4 t2v = cross3(vh, t1v);
    float r = sqrtf(r1);
    float phi = (float)((2.0 * (double)ORC_PI) * (double)r2);
    float t1 = r * mt_xm_cosf(phi);
    float t2 = r * mt_xm_sinf(phi);
    float s = (float)(0.5 * (1.0 + (double)vh.z));
    t2 = (float)((1.0 - (double)s) * sqrt(1.0 - (double)(t1 * t1)) + (double)s * (double)t2);
    float k = sqrtf(fmaxf(0.0f, (1.0f - t1 * t1) - t2 * t2));
    f4 nh = f4_add(f4_add(f4_scale(t1v, t1), f4_scale(t2v, t2)), f4_scale(vh, k));
    return normalize3(f4_make(ax * nh.x, ay * nh.y, fmaxf(0.0f, nh.z), 0.0f));
}

static float ggx_vndf(f4 v, f4 n, float ax, float ay)
{ /* VNDF, Bsdf.cl:297-302 */
    return ((ggx_G1(v, ax, ay) * fmaxf(0.001f, dot3(v, n))) * ggx_D(n, ax, ay)) / fmaxf(0.001f, v.z);
}

static float ggx_pdf(f4 v, f4 ni, float ax, float ay)
{ /* pdfGGXVNDF, Bsdf.cl:305-311 */
    float dv = ggx_vndf(v, ni, ax, ay);
    float vn = dot3(v, ni);
    return dv / (4.0f * vn);
}

/* exported GGX primitives for closed-form tests */
float orc_ggx_D(const float* wh, float ax, float ay) { return ggx_D(f4_make(wh[0], wh[1], wh[2], 0), ax, ay); }
float orc_ggx_lambda(const float* w, float ax, float ay) { return ggx_lambda(f4_make(w[0], w[1], w[2], 0), ax, ay); }
void orc_ggx_sample_vndf(const float* v, float ax, float ay, float u1, float u2, float* out)
{
    f4 h = ggx_sample_vndf(f4_make(v[0], v[1], v[2], 0), ax, ay, u1, u2);
    out[0] = h.x; out[1] = h.y; out[2] = h.z;
}
float orc_ggx_pdf_vndf(const float* v, const float* h, float ax, float ay)
{ return ggx_pdf(f4_make(v[0], v[1], v[2], 0), f4_make(h[0], h[1], h[2], 0), ax, ay); }

/* ---- material record: 4 int flags + 8 float4 (Bsdf.cl:24-42) ---- */
typedef struct {
    const float* m; /* 36 words */
} omat;
static f4 mat_f4(const float* m, int slot) { return f4_load(m + 4 + 4 * slot); }
#define M_DIFF_COLOR 0
#define M_DIFF_PARAM 1
#define M_GLOSSY_COLOR 2
#define M_GLOSSY_PARAM 3

typedef struct {
    oframe F;
    f4 fix, geom;
    float diff_prob, glossy_prob;
    int mat;
    const float* m;
} obsdf;

static void bsdf_unpack(obsdf* B, const float* rec, int mat_id, const float* materials)
{
    B->F.mx = f4_make(rec[0], rec[1], rec[2], 0);
    B->F.my = f4_make(rec[3], rec[4], rec[5], 0);
    B->F.mz = f4_make(rec[6], rec[7], rec[8], 0);
    B->fix = f4_make(rec[9], rec[10], rec[11], 0);
    B->geom = f4_make(rec[12], rec[13], rec[14], 0);
    B->diff_prob = rec[15];
    B->glossy_prob = rec[16];
    B->mat = mat_id;
    B->m = mat_id >= 0 ? materials + 36 * (size_t)mat_id : materials;
}

int orc_bsdf_setup(int32_t n, const float* ray_d, const float* isect_n,
                   const float* isect_ng, const int32_t* mat,
                   const float* materials, float* bsdf, int32_t* mat_id)
{ /* setupBsdf, Bsdf.cl:130-160 */
    for (int32_t i = 0; i < n; i++) {
        float* rec = bsdf + (size_t)ORC_BSDF_WORDS * i;
        for (int k = 0; k < ORC_BSDF_WORDS; k++) rec[k] = 0.0f;
        f4 nn = f4_make(isect_n[4 * i], isect_n[4 * i + 1], isect_n[4 * i + 2], 0);
        f4 d = f4_make(ray_d[4 * i], ray_d[4 * i + 1], ray_d[4 * i + 2], 0);
        f4 ng = f4_make(isect_ng[4 * i], isect_ng[4 * i + 1], isect_ng[4 * i + 2], 0);
        oframe F = get_frame(nn);
        f4 fix = to_local(&F, f4_neg(d));
        rec[0] = F.mx.x; rec[1] = F.mx.y; rec[2] = F.mx.z;
        rec[3] = F.my.x; rec[4] = F.my.y; rec[5] = F.my.z;
        rec[6] = F.mz.x; rec[7] = F.mz.y; rec[8] = F.mz.z;
        rec[9] = fix.x; rec[10] = fix.y; rec[11] = fix.z;
        /* grazing, or no surface (mat -1: a miss; the reference never sets up those) */
        if (fabsf(fix.z) < EPS_COSINE || mat[i] < 0) { mat_id[i] = -1; continue; }
        f4 g = to_local(&F, ng);
        rec[12] = g.x; rec[13] = g.y; rec[14] = g.z;
        mat_id[i] = mat[i];
        const float* m = materials + 36 * (size_t)mat[i];
        /* GetComponentProbabilities, Bsdf.cl:93-128 with AlbedoDiffuse/Glossy :70-80 */
        f4 dc = f4_scale(mat_f4(m, M_DIFF_COLOR), mat_f4(m, M_DIFF_PARAM).x);
        f4 gc = f4_scale(mat_f4(m, M_GLOSSY_COLOR), mat_f4(m, M_GLOSSY_PARAM).x);
        float ad = luminance3(dc.x, dc.y, dc.z);
        float ag = luminance3(gc.x, gc.y, gc.z);
        float tot = ad + ag;
        if (tot < 1e-9f) { rec[15] = 0.0f; rec[16] = 0.0f; }
        else { rec[15] = ad / tot; rec[16] = ag / tot; }
    }
    return ORC_OK;
}

static f4 schlick(float c, f4 f0)
{ /* SchlickColor, Bsdf.cl:185-189 */
    float e = mt_xm_pow5f(1.0f - c);
    return f4_make(f0.x + (1.0f - f0.x) * e, f0.y + (1.0f - f0.y) * e, f0.z + (1.0f - f0.z) * e, 0.0f);
}

static f4 eval_diffuse(const obsdf* B, f4 gen, float* pdf)
{ /* EvaluateDiffuse, Bsdf.cl:452-467 */
    if (B->diff_prob == 0.0f) return f4_make(0, 0, 0, 0);
    if (B->fix.z < EPS_COSINE || gen.z < EPS_COSINE) return f4_make(0, 0, 0, 0);
    if (pdf) *pdf += B->diff_prob * fmaxf(0.0f, (gen.z * 1.0f) / ORC_PI);
    f4 c = mat_f4(B->m, M_DIFF_COLOR);
    return f4_make((c.x * 1.0f) / ORC_PI, (c.y * 1.0f) / ORC_PI, (c.z * 1.0f) / ORC_PI, 1.0f);
}

static f4 glossy_value(const obsdf* B, f4 wi, f4 wo, f4 wh, float ax, float ay)
{ /* shared tail of SampleGlossy / EvaluateGlossy, Bsdf.cl:414-430,489-505 */
    float ci = fabsf(wi.z), co = fabsf(wo.z);
    float fr = (ggx_G(wi, wo, ax, ay) * ggx_D(wh, ax, ay)) / ((4.0f * ci) * co);
    float whdot = fabsf(dot3(wh, wo));
    f4 fc = schlick(whdot, mat_f4(B->m, M_GLOSSY_COLOR));
    if (ci == 0.0f || co == 0.0f) fc = f4_make(0, 0, 0, 0);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) fc = f4_make(0, 0, 0, 0);
    return f4_make(fc.x * fr, fc.y * fr, fc.z * fr, 0.0f);
}

static f4 eval_glossy(const obsdf* B, f4 gen, float* pdf)
{ /* EvaluateGlossy, Bsdf.cl:469-507 */
    if (B->glossy_prob == 0.0f) return f4_make(0, 0, 0, 0);
    if (B->fix.z < EPS_COSINE || gen.z < EPS_COSINE) return f4_make(0, 0, 0, 0);
    f4 gp = mat_f4(B->m, M_GLOSSY_PARAM);
    float ax = gp.y, ay = gp.z;
    f4 wi = B->fix, wo = gen;
    f4 wh = normalize3(f4_add(wi, wo)); /* halfVector, :336-341 */
    *pdf += B->glossy_prob * ggx_pdf(wo, wh, ax, ay);
    return glossy_value(B, wi, wo, wh, ax, ay);
}

static f4 sample_diffuse(const obsdf* B, float s0, float s1, f4* gen, float* pdf)
{ /* SampleDiffuse, Bsdf.cl:380-394 + SampleCosHemisphereW, Sampling.cl:129-148 */
    if (B->fix.z < EPS_COSINE) return f4_make(0, 0, 0, 0);
    float term1 = (2.0f * ORC_PI) * s0;
    float term2 = sqrtf(1.0f - s1);
    *gen = f4_make(mt_xm_cosf(term1) * term2, mt_xm_sinf(term1) * term2, sqrtf(s1), 0.0f);
    float updf = (gen->z * 1.0f) / ORC_PI;
    *pdf += updf * B->diff_prob;
    f4 c = mat_f4(B->m, M_DIFF_COLOR);
    return f4_make((c.x * 1.0f) / ORC_PI, (c.y * 1.0f) / ORC_PI, (c.z * 1.0f) / ORC_PI, 1.0f);
}

static f4 sample_glossy(const obsdf* B, float s0, float s1, f4* gen, float* pdf)
{ /* SampleGlossy, Bsdf.cl:396-432 */
    if (B->fix.z < EPS_COSINE) return f4_make(0, 0, 0, 0);
    f4 gp = mat_f4(B->m, M_GLOSSY_PARAM);
    float ax = gp.y, ay = gp.z;
    f4 wh = ggx_sample_vndf(B->fix, ax, ay, s0, s1);
    float k = 2.0f * dot3(B->fix, wh);
    *gen = f4_add(f4_neg(B->fix), f4_scale(wh, k)); /* reflectFromN, :325-328 */
    gen->w = 0.0f;
    f4 wi = B->fix, wo = *gen;
    *pdf += B->glossy_prob * ggx_pdf(wo, wh, ax, ay);
    return glossy_value(B, wi, wo, wh, ax, ay);
}

int orc_bsdf_sample(int32_t n, const float* bsdf, const int32_t* mat_id,
                    const float* materials, const float* u, float* f,
                    float* dir, float* pdf, float* cos_out)
{ /* SampleBsdf, Bsdf.cl:555-606 */
    for (int32_t i = 0; i < n; i++) {
        obsdf B;
        bsdf_unpack(&B, bsdf + (size_t)ORC_BSDF_WORDS * i, mat_id[i], materials);
        float p = 0.0f, c = 0.0f;
        f4 gen = f4_make(0, 0, 0, 0);
        f4 res = f4_make(0, 0, 0, 1);
        f4 wd = f4_make(0, 0, 0, 0);
        int zero = 0;
        if (u[3 * i + 2] < B.diff_prob) {
            res = f4_add(res, sample_diffuse(&B, u[3 * i], u[3 * i + 1], &gen, &p));
            if (res.x == 0 && res.y == 0 && res.z == 0 && res.w == 0) zero = 1;
            else res = f4_add(res, eval_glossy(&B, gen, &p));
        } else {
            res = f4_add(res, sample_glossy(&B, u[3 * i], u[3 * i + 1], &gen, &p));
            if (res.x == 0 && res.y == 0 && res.z == 0 && res.w == 0) zero = 1;
            else res = f4_add(res, eval_diffuse(&B, gen, &p));
        }
        if (!zero) {
            c = fabsf(gen.z);
            if (c < EPS_COSINE) zero = 1;
            else wd = to_world(&B.F, gen);
        }
        if (zero) res = f4_make(0, 0, 0, 0);
        f[3 * i] = res.x; f[3 * i + 1] = res.y; f[3 * i + 2] = res.z;
        dir[3 * i] = wd.x; dir[3 * i + 1] = wd.y; dir[3 * i + 2] = wd.z;
        pdf[i] = p;
        cos_out[i] = c;
    }
    return ORC_OK;
}

int orc_bsdf_eval(int32_t n, const float* bsdf, const int32_t* mat_id,
                  const float* materials, const float* dir, float* f,
                  float* pdf, float* cos_out)
{ /* EvaluateBsdf, Bsdf.cl:525-545 (pdf and cos start at 0 as in DirectLight, Render.cl:440) */
    for (int32_t i = 0; i < n; i++) {
        obsdf B;
        bsdf_unpack(&B, bsdf + (size_t)ORC_BSDF_WORDS * i, mat_id[i], materials);
        float p = 0.0f, c = 0.0f;
        f4 res = f4_make(0, 0, 0, 1);
        f4 gen = to_local(&B.F, f4_make(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], 0));
        if (!(gen.z * B.fix.z < 0.0f)) {
            c = gen.z;
            res = f4_add(res, eval_diffuse(&B, gen, &p));
            res = f4_add(res, eval_glossy(&B, gen, &p));
        }
        f[3 * i] = res.x; f[3 * i + 1] = res.y; f[3 * i + 2] = res.z;
        pdf[i] = p;
        cos_out[i] = c;
    }
    return ORC_OK;
}

/* ---- RNG: Sampling.cl:15-59 ---- */
static int32_t asr(int32_t x, int s) { return (int32_t)(x >> s); } /* arithmetic shift (gcc) */

uint32_t orc_wang_hash(int32_t seed)
{
    int32_t s = (seed ^ 61) ^ asr(seed, 16);
    s = (int32_t)((uint32_t)s * 9u);
    s = s ^ asr(s, 4);
    s = (int32_t)((uint32_t)s * 0x27d4eb2du);
    s = s ^ asr(s, 15);
    return (uint32_t)s;
}

float orc_get_random(int32_t* st)
{
    st[0] = (int32_t)orc_wang_hash(st[0]);
    st[1] = (int32_t)orc_wang_hash(st[1]);
    uint32_t ires = ((uint32_t)st[0] << 16) + (uint32_t)st[1];
    union { float f; uint32_t u; } r;
    r.u = (ires & 0x007fffffu) | 0x40000000u;
    return (r.f - 2.0f) / 2.0f;
}
