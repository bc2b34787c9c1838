/*
 * Vector-ALU issue-rate calibration (measurement aid, built as its own
 * library, libmt_calib.so; the path-tracing library does not link it).
 *
 * The path-tracing kernels are bound by instruction issue, not by HBM, so
 * their roofline is VALU instructions per second.  The peak of that roofline
 * depends on how many wave64 VALU instructions a SIMD retires per cycle for
 * the instruction mix in question, which the published FLOP figures (counted
 * with packed FP32) do not state for unpacked code.  These kernels measure it:
 * every lane runs KIND's fixed mix of independent chains, so the issue rate is
 * limited by the SIMDs and not by latency (8 chains per lane, many waves per
 * SIMD).
 *
 *   kind 0  8 v_mul_f32 + 8 v_add_f32 per iteration (16)
 *   kind 1  4 v_mul_f32 + 8 v_add_f32 + 4 v_rcp_f32 (16, a quarter of them
 *           transcendental)
 *   kind 2  v_min_f32/v_max_f32, the slab test's instructions: 16 written,
 *           plus 8 v_max_f32 x, x, x that quiet signalling NaNs before
 *           fminf/fmaxf (24)
 *   kind 3  8 v_add_u32 + 8 v_xor_b32 (16)
 *   kind 4  8 v_cmp_f32 + 8 v_cndmask_b32 + 8 v_add_f32, a select per
 *           chain as the traversal's branch-free code makes them (24)
 *   kind 5  16 v_fma_f32 (16)
 *   kind 6  8 correctly rounded divisions, 11 instructions each: 2
 *           v_div_scale_f32, v_rcp_f32, 5 FMA, v_mul_f32, v_div_fmas_f32,
 *           v_div_fixup_f32 (88)
 *   kind 7  8 correctly rounded square roots, 16 instructions each
 *           (v_sqrt_f32, neighbour checks by FMA, compares and selects),
 *           + 8 v_add_f32 (136)
 *
 * A launch issues that count x iters x (blocks x 256 / 64) wave64 VALU
 * instructions in its loop, plus a few per wave before and after it.
 */
#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

constexpr int CALIB_BLOCK = 256;

template <int KIND>
__global__ __launch_bounds__(CALIB_BLOCK) void k_valu_issue(float* out, int iters, float seed)
{
    const int gid = blockIdx.x * CALIB_BLOCK + threadIdx.x;
    float a0 = seed + 0.001f * (float)threadIdx.x, a1 = a0 + 0.1f, a2 = a0 + 0.2f, a3 = a0 + 0.3f;
    float a4 = a0 + 0.4f, a5 = a0 + 0.5f, a6 = a0 + 0.6f, a7 = a0 + 0.7f;
    for (int i = 0; i < iters; ++i) {
        if (KIND == 0) {
            a0 = a0 * 0.9999f + 0.5f; a1 = a1 * 0.9998f + 0.5f;
            a2 = a2 * 0.9997f + 0.5f; a3 = a3 * 0.9996f + 0.5f;
            a4 = a4 * 0.9995f + 0.5f; a5 = a5 * 0.9994f + 0.5f;
            a6 = a6 * 0.9993f + 0.5f; a7 = a7 * 0.9992f + 0.5f;
        } else if (KIND == 1) {
            a0 = a0 * 0.9999f + 0.5f; a1 = a1 * 0.9998f + 0.5f;
            a2 = a2 * 0.9997f + 0.5f; a3 = a3 * 0.9996f + 0.5f;
            a4 = __builtin_amdgcn_rcpf(a4 + 1.5f); a5 = __builtin_amdgcn_rcpf(a5 + 1.5f);
            a6 = __builtin_amdgcn_rcpf(a6 + 1.5f); a7 = __builtin_amdgcn_rcpf(a7 + 1.5f);
        } else if (KIND == 3) {
            a0 = __uint_as_float((__float_as_uint(a0) + 0x9e37u) ^ 0x55u);
            a1 = __uint_as_float((__float_as_uint(a1) + 0x9e39u) ^ 0x56u);
            a2 = __uint_as_float((__float_as_uint(a2) + 0x9e3bu) ^ 0x57u);
            a3 = __uint_as_float((__float_as_uint(a3) + 0x9e3du) ^ 0x58u);
            a4 = __uint_as_float((__float_as_uint(a4) + 0x9e3fu) ^ 0x59u);
            a5 = __uint_as_float((__float_as_uint(a5) + 0x9e41u) ^ 0x5au);
            a6 = __uint_as_float((__float_as_uint(a6) + 0x9e43u) ^ 0x5bu);
            a7 = __uint_as_float((__float_as_uint(a7) + 0x9e45u) ^ 0x5cu);
        } else if (KIND == 4) {
            a0 = a0 < a4 ? a4 : a0 + 1.0f; a1 = a1 < a5 ? a5 : a1 + 1.0f;
            a2 = a2 < a6 ? a6 : a2 + 1.0f; a3 = a3 < a7 ? a7 : a3 + 1.0f;
            a4 = a4 > a0 ? a0 : a4 - 1.0f; a5 = a5 > a1 ? a1 : a5 - 1.0f;
            a6 = a6 > a2 ? a2 : a6 - 1.0f; a7 = a7 > a3 ? a3 : a7 - 1.0f;
        } else if (KIND == 5) {
            a0 = fmaf(fmaf(a0, 0.9999f, 0.5f), 0.9999f, 0.5f); a1 = fmaf(fmaf(a1, 0.9998f, 0.5f), 0.9998f, 0.5f);
            a2 = fmaf(fmaf(a2, 0.9997f, 0.5f), 0.9997f, 0.5f); a3 = fmaf(fmaf(a3, 0.9996f, 0.5f), 0.9996f, 0.5f);
            a4 = fmaf(fmaf(a4, 0.9995f, 0.5f), 0.9995f, 0.5f); a5 = fmaf(fmaf(a5, 0.9994f, 0.5f), 0.9994f, 0.5f);
            a6 = fmaf(fmaf(a6, 0.9993f, 0.5f), 0.9993f, 0.5f); a7 = fmaf(fmaf(a7, 0.9992f, 0.5f), 0.9992f, 0.5f);
        } else if (KIND == 6) {
            a0 = 2.0f / a0; a1 = 3.0f / a1; a2 = 5.0f / a2; a3 = 6.0f / a3;
            a4 = 7.0f / a4; a5 = 8.0f / a5; a6 = 10.0f / a6; a7 = 11.0f / a7;
        } else if (KIND == 7) {
            a0 = sqrtf(a0 + 1.0f); a1 = sqrtf(a1 + 1.0f); a2 = sqrtf(a2 + 1.0f); a3 = sqrtf(a3 + 1.0f);
            a4 = sqrtf(a4 + 1.0f); a5 = sqrtf(a5 + 1.0f); a6 = sqrtf(a6 + 1.0f); a7 = sqrtf(a7 + 1.0f);
        } else {
            a0 = fmaxf(fminf(a0, a4), -1.0f); a1 = fmaxf(fminf(a1, a5), -1.0f);
            a2 = fmaxf(fminf(a2, a6), -1.0f); a3 = fmaxf(fminf(a3, a7), -1.0f);
            a4 = fminf(fmaxf(a4, a0), 2.0f); a5 = fminf(fmaxf(a5, a1), 2.0f);
            a6 = fminf(fmaxf(a6, a2), 2.0f); a7 = fminf(fmaxf(a7, a3), 2.0f);
        }
    }
    out[gid] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

} // namespace

extern "C" {

/* Launches `blocks` work-groups of 256 lanes of mix `kind` (0..7) running
   `iters` iterations on `stream`; `out` holds blocks x 256 floats.  Returns 0,
   or -1 for an argument out of range, or the HIP error of the launch. */
int mt_calib_valu(int kind, int blocks, int iters, float* out, void* stream)
{
    if (kind < 0 || kind > 7 || blocks <= 0 || blocks > (1 << 20) || iters <= 0 || out == nullptr) return -1;
    const hipStream_t s = (hipStream_t)stream;
    const dim3 g(blocks), b(CALIB_BLOCK);
    switch (kind) {
    case 0: hipLaunchKernelGGL(k_valu_issue<0>, g, b, 0, s, out, iters, 1.0f); break;
    case 1: hipLaunchKernelGGL(k_valu_issue<1>, g, b, 0, s, out, iters, 1.0f); break;
    case 2: hipLaunchKernelGGL(k_valu_issue<2>, g, b, 0, s, out, iters, 1.0f); break;
    case 3: hipLaunchKernelGGL(k_valu_issue<3>, g, b, 0, s, out, iters, 1.0f); break;
    case 4: hipLaunchKernelGGL(k_valu_issue<4>, g, b, 0, s, out, iters, 1.0f); break;
    case 5: hipLaunchKernelGGL(k_valu_issue<5>, g, b, 0, s, out, iters, 1.0f); break;
    case 6: hipLaunchKernelGGL(k_valu_issue<6>, g, b, 0, s, out, iters, 1.0f); break;
    default: hipLaunchKernelGGL(k_valu_issue<7>, g, b, 0, s, out, iters, 1.0f); break;
    }
    return (int)hipGetLastError();
}

} // extern "C"
