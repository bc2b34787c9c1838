/*
 * Transcendental functions of the path-tracing core with one definition for
 * the GPU kernels and the host: sin, cos, acos, atan2 and the fifth power of
 * SchlickColor (pow(1 - cos, 5), Bsdf.cl:185-189).  The reference calls the
 * OpenCL builtins, whose last bits depend on the vendor library; here every
 * result is a fixed sequence of IEEE double operations (+, -, *, /, sqrt,
 * floor; no fused multiply-add) rounded once to float, so device and host
 * agree bit for bit and the results are within one float ulp of the exact
 * values (nearly always correctly rounded).
 *
 * Methods (textbook; coefficients are exact Taylor coefficients 1/n!
 * and 1/n rounded to double):
 *   sin, cos   Cody-Waite reduction by pi/2 with a two-part constant (pi/2
 *              = HI + LO, HI holding 33 significant bits so k * HI is exact
 *              for |k| < 2^20), then the Taylor series on [-pi/4, pi/4] to
 *              degree 15 (sin) and 16 (cos); valid for |x| < 2^20.
 *   atan2      ratio min/max in [0, 1]; above 2 - sqrt(3) the argument is
 *              moved by the addition formula atan z = pi/6 +
 *              atan((z sqrt3 - 1)/(z + sqrt3)); the Taylor series of atan to
 *              degree 27 on |z| <= 2 - sqrt(3); quadrants and signed zeros
 *              as C99 atan2.
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

#define MT_XM_PI 3.141592653589793
#define MT_XM_PIO2 1.5707963267948966
#define MT_XM_PIO2_HI 1.5707963267341256 /* pi/2 truncated to 33 bits */
#define MT_XM_PIO2_LO 6.077100506506192e-11 /* pi/2 - HI */
#define MT_XM_PI6 0.5235987755982989
#define MT_XM_SQRT3 1.7320508075688772
#define MT_XM_TAN_PI12 0.2679491924311228 /* 2 - sqrt(3) */
#define MT_XM_2OPI 0.6366197723675814

/* sin(r), cos(r) for |r| <= pi/4: Taylor series, Horner in r^2 */
MT_XM_FN double mt_xm_sin_poly(double r)
{
    const double z = r * r;
    double p = 7.647163731819816e-13;             /* 1/15! */
    p = p * z - 1.6059043836821613e-10;           /* 1/13! */
    p = p * z + 2.505210838544172e-08;            /* 1/11! */
    p = p * z - 2.7557319223985893e-06;           /* 1/9! */
    p = p * z + 0.0001984126984126984;            /* 1/7! */
    p = p * z - 0.008333333333333333;             /* 1/5! */
    p = p * z + 0.16666666666666666;              /* 1/3! */
    return r - (r * z) * p;
}

MT_XM_FN double mt_xm_cos_poly(double r)
{
    const double z = r * r;
    double p = 4.779477332387385e-14;             /* 1/16! */
    p = p * z - 1.1470745597729725e-11;           /* 1/14! */
    p = p * z + 2.08767569878681e-09;             /* 1/12! */
    p = p * z - 2.755731922398589e-07;            /* 1/10! */
    p = p * z + 2.48015873015873e-05;             /* 1/8! */
    p = p * z - 0.001388888888888889;             /* 1/6! */
    p = p * z + 0.041666666666666664;             /* 1/4! */
    return (1.0 - 0.5 * z) + (z * z) * p;
}

/* quadrant q = k mod 4 and reduced argument r = x - k pi/2 */
MT_XM_FN double mt_xm_reduce(double x, int* q)
{
    const double k = floor(x * MT_XM_2OPI + 0.5);
    *q = (int)(k - 4.0 * floor(k * 0.25));
    return (x - k * MT_XM_PIO2_HI) - k * MT_XM_PIO2_LO;
}

MT_XM_FN float mt_xm_sinf(float xf)
{
    int q;
    const double r = mt_xm_reduce((double)xf, &q);
    const double v = (q & 1) ? mt_xm_cos_poly(r) : mt_xm_sin_poly(r);
    return (float)((q & 2) ? -v : v);
}

MT_XM_FN float mt_xm_cosf(float xf)
{
    int q;
    const double r = mt_xm_reduce((double)xf, &q);
    const double v = (q & 1) ? mt_xm_sin_poly(r) : mt_xm_cos_poly(r);
    return (float)(((q + 1) & 2) ? -v : v);
}

/* atan(z) for |z| <= 2 - sqrt(3): Taylor series to z^27 */
MT_XM_FN double mt_xm_atan_poly(double t)
{
    const double z = t * t;
    double p = 1.0 / 27.0;
    p = p * z - 1.0 / 25.0;
    p = p * z + 1.0 / 23.0;
    p = p * z - 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z - 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z - 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z - 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z - 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    return t - (t * z) * p;
}

MT_XM_FN double mt_xm_atan2(double y, double x)
{
    if (x != x || y != y) return x + y; /* NaN */
    const double ax = fabs(x), ay = fabs(y);
    const int xneg = signbit(x) != 0, yneg = signbit(y) != 0;
    double a;
    if (ay == 0.0 && ax == 0.0) {
        a = xneg ? MT_XM_PI : 0.0;
    } else {
        const int swap = ay > ax;
        double z;
        if (ax == ay) z = 1.0; /* also both infinite */
        else z = swap ? ax / ay : ay / ax;
        if (z > MT_XM_TAN_PI12) a = MT_XM_PI6 + mt_xm_atan_poly((z * MT_XM_SQRT3 - 1.0) / (z + MT_XM_SQRT3));
        else a = mt_xm_atan_poly(z);
        if (swap) a = MT_XM_PIO2 - a;
        if (xneg) a = MT_XM_PI - a;
    }
    return yneg ? -a : a;
}

MT_XM_FN float mt_xm_atan2f(float y, float x) { return (float)mt_xm_atan2((double)y, (double)x); }

MT_XM_FN float mt_xm_acosf(float yf)
{
    const double y = (double)yf;
    return (float)mt_xm_atan2(sqrt((1.0 - y) * (1.0 + y)), y);
}

/* x^5 (pow(x, 5.f) of SchlickColor) */
MT_XM_FN float mt_xm_pow5f(float xf)
{
    const double x = (double)xf, x2 = x * x;
    return (float)((x2 * x2) * x);
}

#endif
