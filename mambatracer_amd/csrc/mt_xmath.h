/*
 * Transcendental functions of the path-tracing core with one definition for
 * the GPU kernels and the host: sin, cos, acos, atan2 and the fifth power of
 * SchlickColor (pow(1 - cos, 5), Bsdf.cl:185-189).  The reference calls the
 * OpenCL builtins, whose last bits depend on the vendor library; here every
 * result is a fixed sequence of IEEE single-precision operations (+, -, *,
 * /, sqrt, floor; no fused multiply-add), so device and host agree bit for
 * bit; results are within 2 ulp of the exact values (most are correctly
 * rounded).
 *
 * Methods (textbook; coefficients are the Taylor coefficients 1/n! and 1/n
 * rounded to float; each series is cut where its remainder is below 1e-8
 * relative):
 *   sin, cos   Cody-Waite reduction by pi/2 with a three-part constant
 *              (pi/2 = C1 + C2 + C3, C1 and C2 holding 12 significant bits,
 *              so k * C1 and k * C2 are exact for |k| < 2^12), then the
 *              Taylor series on [-pi/4, pi/4] to degree 9 (sin) and 10
 *              (cos); valid for |x| < 6000.
 *   atan2      ratio min/max in [0, 1]; above 2 - sqrt(3) the argument is
 *              moved by the addition formula atan z = pi/6 +
 *              atan((z sqrt3 - 1)/(z + sqrt3)); the Taylor series of atan to
 *              degree 13 on |z| <= 2 - sqrt(3); quadrants and signed zeros
 *              as C99 atan2; pi/2 and pi are applied as a float head plus a
 *              tail.
 *   acos       atan2(sqrt((1 - y)(1 + y)), y).
 *
 * Usable from C and from HIP: the includer defines MT_XM_FN, the function
 * qualifiers (HIP: `__host__ __device__ static inline`; C: `static inline`).
 */
#ifndef MT_XMATH_H
#define MT_XMATH_H

#ifndef MT_XM_FN
#error "define MT_XM_FN (function qualifiers) before including mt_xmath.h"
#endif

#define MT_XM_C1 1.5703125f               /* pi/2 = C1 + C2 + C3 */
#define MT_XM_C2 0.0004837512969970703f
#define MT_XM_C3 7.549790126404332e-08f
#define MT_XM_2OPI 0.6366197466850281f     /* 2/pi */
#define MT_XM_PIO2_HI 1.5707963705062866f  /* float(pi/2) */
#define MT_XM_PIO2_LO -4.371139000186241e-08f /* pi/2 - float(pi/2) */
#define MT_XM_PI_HI 3.1415927410125732f    /* float(pi) */
#define MT_XM_PI_LO -8.742278000372482e-08f /* pi - float(pi) */
#define MT_XM_PI6 0.5235987901687622f      /* float(pi/6) */
#define MT_XM_PI4 0.7853981852531433f      /* float(pi/4) */
#define MT_XM_3PI4 2.356194496154785f      /* float(3pi/4) */
#define MT_XM_SQRT3 1.7320507764816284f
#define MT_XM_TAN_PI12 0.2679491937160492f /* float(2 - sqrt(3)) */

/* sin(r), cos(r) for |r| <= pi/4: Taylor series, Horner in r^2 */
MT_XM_FN float mt_xm_sin_poly(float r)
{
    const float z = r * r;
    float p = 2.7557318844628753e-06f;     /* 1/9! */
    p = p * z - 0.00019841270113829523f;   /* 1/7! */
    p = p * z + 0.008333333767950535f;     /* 1/5! */
    p = p * z - 0.1666666716337204f;       /* 1/3! */
    return r + (r * z) * p;
}

MT_XM_FN float mt_xm_cos_poly(float r)
{
    const float z = r * r;
    float p = -2.755731998149713e-07f;     /* 1/10! */
    p = p * z + 2.4801587642286904e-05f;   /* 1/8! */
    p = p * z - 0.0013888889225199819f;    /* 1/6! */
    p = p * z + 0.0416666679084301f;       /* 1/4! */
    return (1.0f - 0.5f * z) + (z * z) * p;
}

/* quadrant q = k mod 4 and reduced argument r = x - k pi/2 */
MT_XM_FN float mt_xm_reduce(float x, int* q)
{
    const float k = floorf(x * MT_XM_2OPI + 0.5f);
    *q = (int)(k - 4.0f * floorf(k * 0.25f));
    return ((x - k * MT_XM_C1) - k * MT_XM_C2) - k * MT_XM_C3;
}

MT_XM_FN float mt_xm_sinf(float x)
{
    int q;
    const float r = mt_xm_reduce(x, &q);
    const float v = (q & 1) ? mt_xm_cos_poly(r) : mt_xm_sin_poly(r);
    return (q & 2) ? -v : v;
}

MT_XM_FN float mt_xm_cosf(float x)
{
    int q;
    const float r = mt_xm_reduce(x, &q);
    const float v = (q & 1) ? mt_xm_sin_poly(r) : mt_xm_cos_poly(r);
    return ((q + 1) & 2) ? -v : v;
}

/* atan(t) for |t| <= 2 - sqrt(3): Taylor series to t^13 */
MT_XM_FN float mt_xm_atan_poly(float t)
{
    const float z = t * t;
    float p = 0.07692307978868484f;        /* 1/13 */
    p = p * z - 0.09090909361839294f;      /* 1/11 */
    p = p * z + 0.1111111119389534f;       /* 1/9 */
    p = p * z - 0.1428571492433548f;       /* 1/7 */
    p = p * z + 0.20000000298023224f;      /* 1/5 */
    p = p * z - 0.3333333432674408f;       /* 1/3 */
    return t + (t * z) * p;
}

/* atan(z) for z in [0, 1].  Written with selects rather than branches (the
   same operations on every input, so the same results): lanes of one GPU
   wave take both sides, and a branch would evaluate the series twice. */
MT_XM_FN float mt_xm_atan01(float z)
{
    const int red = z > MT_XM_TAN_PI12;
    const float t = red ? (z * MT_XM_SQRT3 - 1.0f) / (z + MT_XM_SQRT3) : z;
    const float p = mt_xm_atan_poly(t);
    return red ? MT_XM_PI6 + p : p;
}

/* One ratio min/max and one series for every input; the octant and the
   special cases are selected afterwards (results as the case analysis:
   NaN -> x + y; both zero -> 0 or pi; equal magnitudes -> pi/4 or 3pi/4;
   |y| > |x| -> pi/2 -+ atan(|x|/|y|); else atan(|y|/|x|) or pi - it). */
MT_XM_FN float mt_xm_atan2f(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const int xneg = signbit(x) != 0, yneg = signbit(y) != 0;
    const int yax = ay > ax; /* angle from the y axis */
    const float t = mt_xm_atan01((yax ? ax : ay) / (yax ? ay : ax));
    float a = yax ? (xneg ? (MT_XM_PIO2_HI + t) + MT_XM_PIO2_LO : (MT_XM_PIO2_HI - t) + MT_XM_PIO2_LO)
                  : (xneg ? (MT_XM_PI_HI - t) + MT_XM_PI_LO : t);
    a = ax == ay ? (xneg ? MT_XM_3PI4 : MT_XM_PI4) : a; /* also both infinite */
    a = (ay == 0.0f && ax == 0.0f) ? (xneg ? MT_XM_PI_HI : 0.0f) : a;
    a = yneg ? -a : a;
    return (x != x || y != y) ? x + y : a;
}

MT_XM_FN float mt_xm_acosf(float y) { return mt_xm_atan2f(sqrtf((1.0f - y) * (1.0f + y)), y); }

/* x^5 (pow(x, 5.f) of SchlickColor) */
MT_XM_FN float mt_xm_pow5f(float x)
{
    const float x2 = x * x;
    return (x2 * x2) * x;
}

#endif
