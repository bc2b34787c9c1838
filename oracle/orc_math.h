/*
 * Oracle (test infrastructure only): scalar float helpers with a pinned
 * evaluation order.  The reference's OpenCL C leaves the order of dot(),
 * cross() and normalize() to the driver; the oracle fixes one order and the
 * HIP kernels follow the same order with FMA contraction off, so both sides
 * round identically.
 *
 *   dot3(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross(a,b)  = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   length(v)   = sqrtf(dot3(v,v))
 *   normalize(v)= v * (1.0f / length(v))
 *
 * The reference's float4 values carry .w = 0 (src/cl/data/CPoint3.java:21),
 * so the fourth product of OpenCL's float4 dot() adds +0 and is dropped.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <math.h>
#include <stddef.h>
#include <stdint.h>

/* sin / cos / acos / atan2 / x^5: the one definition the GPU kernels also
 * compile (mambatracer_amd/csrc/mt_xmath.h), so both round identically;
 * pinned on its own against float64 numpy in tests/test_xmath.py. */
#define MT_XM_FN static inline
#include "mt_xmath.h"

typedef struct { float x, y, z, w; } f4;

static inline f4 f4_make(float x, float y, float z, float w) { f4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
static inline f4 f4_load(const float* p) { return f4_make(p[0], p[1], p[2], p[3]); }
static inline f4 f4_add(f4 a, f4 b) { return f4_make(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline f4 f4_sub(f4 a, f4 b) { return f4_make(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static inline f4 f4_scale(f4 a, float s) { return f4_make(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline f4 f4_mul(f4 a, f4 b) { return f4_make(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
static inline f4 f4_neg(f4 a) { return f4_make(-a.x, -a.y, -a.z, -a.w); }
static inline float dot3(f4 a, f4 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline f4 cross3(f4 a, f4 b)
{
    return f4_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f);
}
static inline float length3(f4 v) { return sqrtf(dot3(v, v)); }
static inline f4 normalize3(f4 v)
{
    float inv = 1.0f / sqrtf(dot3(v, v));
    return f4_make(v.x * inv, v.y * inv, v.z * inv, 0.0f);
}

/* RGBSpace.cl:28-32 */
static inline float luminance3(float r, float g, float b)
{
    return (0.2989f * r + 0.5866f * g) + 0.1145f * b;
}

/* Util.cl:7-8: M_PI is the float literal 3.14159265359f and M_1_PI expands
 * unparenthesised to 1.f/M_PI, so `x * M_1_PI` evaluates as (x * 1.f) / M_PI. */
#define ORC_PI 3.14159265359f

#endif
