// Device-side float helpers with a pinned evaluation order.
//
// The reference's OpenCL C leaves dot()/cross()/normalize() ordering and FMA
// contraction to the driver.  These helpers fix one order, the translation
// units are compiled with -ffp-contract=off, and division / sqrt are the
// correctly rounded IEEE operations, so results round exactly as the CPU
// restatement used by the tests does:
//   dot3(a,b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross3(a,b)  = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize3(v)= v * (1.0f / sqrtf(dot3(v,v)))
// Reference float4 values carry w = 0, so OpenCL's fourth dot term (+0) is
// dropped.
#pragma once

#include <hip/hip_runtime.h>

// sin / cos / acos / atan2 / x^5 shared with the host (one definition, same bits)
#define MT_XM_FN __host__ __device__ static inline
#include "mt_xmath.h"

#define MT_PI 3.14159265359f   // Util.cl:7 (M_PI redefined as this float)
#define MT_EPS_COSINE 1e-6f    // Util.cl:5
#define MT_EPS_RAY 1e-3f       // Util.cl:6
#define MT_FLOATMAX 3.402823e+38f  // Util.cl:1

struct v3 {
    float x, y, z;
};

__device__ __forceinline__ v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 scl3(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 neg3(v3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ v3 cross3(v3 a, v3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len3(v3 v) { return sqrtf(dot3(v, v)); }
__device__ __forceinline__ v3 norm3(v3 v)
{
    float inv = 1.0f / sqrtf(dot3(v, v));
    return mk3(v.x * inv, v.y * inv, v.z * inv);
}
__device__ __forceinline__ v3 xyz(float4 a) { return mk3(a.x, a.y, a.z); }

// RGBSpace.cl:28-32
__device__ __forceinline__ float luminance3(float r, float g, float b)
{
    return (0.2989f * r + 0.5866f * g) + 0.1145f * b;
}
