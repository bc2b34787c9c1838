/* Oracle (test infrastructure only): CImage.processImage and PrefixSumFloat.
 *
 * Restates averageAccum / updateFrameImage (src/cl/kernel/Image.cl:29-118,
 * RGBSpace.cl:23-106, getIntARGB Util.cl:469-480, check Util.cl:381-387)
 * pixel by pixel in float, and the fixed summation order the device uses
 * for the log-luminance totals and the float scan (2048-element tiles =
 * 256 lanes x 8 rows; a wave's 64 lane values reduced by shuffle-down steps
 * 32, 16, ..., 1 where a lane past the end reads its own value; the four
 * waves added in order; tile partials reduced the same way by one group).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define OT 256
#define OROWS 8
#define OTILE (OT * OROWS)

static float lum_ref(float r, float g, float b) { return (0.2989f * r + 0.5866f * g) + 0.1145f * b; }

static float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

static int32_t argb(float r, float g, float b, float a)
{
    int ri = (int)(clamp01(r) * 255.0f), gi = (int)(clamp01(g) * 255.0f);
    int bi = (int)(clamp01(b) * 255.0f), ai = (int)(clamp01(a) * 255.0f);
    return (int32_t)(((uint32_t)ai << 24) | ((uint32_t)ri << 16) | ((uint32_t)gi << 8) | (uint32_t)bi);
}

/* shuffle-down reduction of 64 lane values; returns lane 0 */
static float wave_sum_f(float* s)
{
    for (int o = 32; o > 0; o >>= 1) {
        float t[64];
        for (int l = 0; l < 64; l++) t[l] = s[l] + (l + o < 64 ? s[l + o] : s[l]);
        memcpy(s, t, sizeof(t));
    }
    return s[0];
}

static int32_t wave_sum_i(int32_t* s)
{
    for (int o = 32; o > 0; o >>= 1) {
        int32_t t[64];
        for (int l = 0; l < 64; l++) t[l] = (int32_t)((uint32_t)s[l] + (uint32_t)(l + o < 64 ? s[l + o] : s[l]));
        memcpy(s, t, sizeof(t));
    }
    return s[0];
}

/* tile_sum: lane t sums rows k = 0..7 (element base + k*256 + t) in order */
static float tile_sum_f(const float* x, int64_t base, int64_t n)
{
    float lane[OT];
    for (int t = 0; t < OT; t++) {
        float s = 0.0f;
        for (int k = 0; k < OROWS; k++) {
            int64_t i = base + (int64_t)k * OT + t;
            s = s + (i < n ? x[i] : 0.0f);
        }
        lane[t] = s;
    }
    float w[4];
    for (int q = 0; q < 4; q++) w[q] = wave_sum_f(lane + 64 * q);
    return ((w[0] + w[1]) + w[2]) + w[3];
}

static int32_t tile_sum_i(const int32_t* x, int64_t base, int64_t n)
{
    int32_t lane[OT];
    for (int t = 0; t < OT; t++) {
        int32_t s = 0;
        for (int k = 0; k < OROWS; k++) {
            int64_t i = base + (int64_t)k * OT + t;
            s += (i < n ? x[i] : 0);
        }
        lane[t] = s;
    }
    int32_t w[4];
    for (int q = 0; q < 4; q++) w[q] = wave_sum_i(lane + 64 * q);
    return ((w[0] + w[1]) + w[2]) + w[3];
}

int32_t orc_image_process(const float* accum, float frame_count, int32_t n, float exposure, float gamma,
                          int32_t update_totals, float* frame, int32_t* argb_io, float* total_log,
                          int32_t* total_count)
{
    if (n <= 0) return -1;
    float* loglw = (float*)malloc(sizeof(float) * (size_t)n);
    int32_t* cnt = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!loglw || !cnt) {
        free(loglw);
        free(cnt);
        return -3;
    }
    for (int32_t i = 0; i < n; i++) { /* averageAccum, Image.cl:29-65 */
        const float* a = accum + 4 * (size_t)i;
        float* f = frame + 4 * (size_t)i;
        f[0] = a[0] / frame_count;
        f[1] = a[1] / frame_count;
        f[2] = a[2] / frame_count;
        f[3] = 1.0f;
        float lw = lum_ref(f[0], f[1], f[2]);
        loglw[i] = lw > 0.0f ? logf(0.01f + lw) : 0.0f;
        cnt[i] = lw > 0.0f ? 1 : 0;
    }
    if (update_totals) {
        int64_t ntiles = ((int64_t)n + OTILE - 1) / OTILE;
        float* fp = (float*)malloc(sizeof(float) * (size_t)ntiles);
        int32_t* cp = (int32_t*)malloc(sizeof(int32_t) * (size_t)ntiles);
        if (!fp || !cp) {
            free(fp);
            free(cp);
            free(loglw);
            free(cnt);
            return -3;
        }
        for (int64_t t = 0; t < ntiles; t++) {
            fp[t] = tile_sum_f(loglw, t * OTILE, n);
            cp[t] = tile_sum_i(cnt, t * OTILE, n);
        }
        float lf[OT];
        int32_t li[OT];
        for (int t = 0; t < OT; t++) {
            float a = 0.0f;
            int32_t b = 0;
            for (int64_t i = t; i < ntiles; i += OT) {
                a = a + fp[i];
                b += cp[i];
            }
            lf[t] = a;
            li[t] = b;
        }
        float wf[4];
        int32_t wi[4];
        for (int q = 0; q < 4; q++) {
            wf[q] = wave_sum_f(lf + 64 * q);
            wi[q] = wave_sum_i(li + 64 * q);
        }
        *total_log = ((wf[0] + wf[1]) + wf[2]) + wf[3];
        *total_count = ((wi[0] + wi[1]) + wi[2]) + wi[3];
        free(fp);
        free(cp);
    }
    float avg = expf(*total_log / (float)*total_count);
    for (int32_t i = 0; i < n; i++) { /* updateFrameImage, Image.cl:68-118 */
        float* c = frame + 4 * (size_t)i;
        float cl = lum_ref(c[0], c[1], c[2]);
        if (!(cl > 0.0f)) continue;
        float scaled = (exposure * cl) / avg;
        float Y = scaled / (1.0f + scaled);
        if (isinf(Y) || isnan(Y) || Y < 0.000001f) Y = 0.0f;
        float X = (0.5893f * c[0] + 0.1789f * c[1]) + 0.1831f * c[2];
        float Yc = (0.2904f * c[0] + 0.6051f * c[1]) + 0.1045f * c[2];
        float Z = (0.0000f * c[0] + 0.0684f * c[1]) + 1.0202f * c[2];
        float sum = (X + Yc) + Z, x = 0.0f, y = 0.0f;
        if (!(sum < 1e-6f)) {
            float s = 1.0f / sum;
            x = X * s;
            y = Yc * s;
        }
        float X2 = y < 1e-6f ? 0.0f : Y * (x / y);
        float Z2 = y < 1e-6f ? 0.0f : (Y * ((1.0f - x) - y)) / y;
        float r = (1.967f * X2 - 0.548f * Y) - 0.297f * Z2;
        float g = (-0.955f * X2 + 1.938f * Y) - 0.027f * Z2;
        float b = (0.064f * X2 - 0.130f * Y) + 0.982f * Z2;
        float ig = 1.0f / gamma;
        c[0] = powf(r, ig);
        c[1] = powf(g, ig);
        c[2] = powf(b, ig);
        c[3] = 1.0f;
        argb_io[i] = argb(c[0], c[1], c[2], 1.0f);
    }
    free(loglw);
    free(cnt);
    return 0;
}

int32_t orc_image_init_argb(void) { return argb(0.0f, 0.0f, 0.0f, 1.0f); }

int32_t orc_argb(float r, float g, float b, float a) { return argb(r, g, b, a); }

/* The device float scan's order: tile partials by tile_sum; their exclusive
 * scan with chunks of ceil(ntiles/256) per lane then chunk offsets in order;
 * within a tile lane t owns elements base + 8t .. 8t+7 (summed in order),
 * lane totals accumulated in lane order from the tile's offset. */
int32_t orc_scan_excl_f32(const float* in, int64_t n, float* out, float* total)
{
    if (n <= 0) {
        *total = 0.0f;
        return 0;
    }
    int64_t ntiles = (n + OTILE - 1) / OTILE;
    float* part = (float*)malloc(sizeof(float) * (size_t)ntiles);
    if (!part) return -3;
    for (int64_t t = 0; t < ntiles; t++) part[t] = tile_sum_f(in, t * OTILE, n);
    int64_t per = (ntiles + OT - 1) / OT;
    float chunk[OT];
    for (int t = 0; t < OT; t++) {
        float run = 0.0f;
        for (int64_t i = t * per; i < (t + 1) * per && i < ntiles; i++) run = run + part[i];
        chunk[t] = run;
    }
    float acc = 0.0f;
    for (int t = 0; t < OT; t++) {
        float v = chunk[t];
        chunk[t] = acc;
        acc = acc + v;
    }
    *total = acc;
    for (int t = 0; t < OT; t++) {
        float a = chunk[t];
        for (int64_t i = t * per; i < (t + 1) * per && i < ntiles; i++) {
            float v = part[i];
            part[i] = a;
            a = a + v;
        }
    }
    for (int64_t tile = 0; tile < ntiles; tile++) {
        int64_t base = tile * OTILE;
        float lane[OT], v[OT][OROWS];
        for (int t = 0; t < OT; t++) {
            float s = 0.0f;
            for (int k = 0; k < OROWS; k++) {
                int64_t i = base + (int64_t)t * OROWS + k;
                v[t][k] = i < n ? in[i] : 0.0f;
                s = s + v[t][k];
            }
            lane[t] = s;
        }
        float a = part[tile];
        for (int t = 0; t < OT; t++) {
            float s = lane[t];
            lane[t] = a;
            a = a + s;
        }
        for (int t = 0; t < OT; t++) {
            float b = lane[t];
            for (int k = 0; k < OROWS; k++) {
                int64_t i = base + (int64_t)t * OROWS + k;
                if (i < n) out[i] = b;
                b = b + v[t][k];
            }
        }
    }
    free(part);
    return 0;
}
