// Context, device buffers and error reporting.
//
// Replaces the un-vendored wrapper.core.OpenCLConfiguration / CMemory layer
// the Java host uses to create buffers, move data and finish the queue
// (e.g. src/cl/device/CDeviceGI.java:150-179 and :428).  One context owns one
// HIP stream on one device, mirroring the reference's one in-order OpenCL
// queue per device.
#include "mt_internal.h"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

__global__ void k_fill_i32(int32_t* dst, int32_t v, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

extern "C" int32_t mt_ctx_create(int32_t device, void* stream, int32_t flags, mt_context** out)
{
    MT_ARG_CHECK(out != nullptr, "mt_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    MT_HIP_CHECK(hipGetDeviceCount(&ndev));
    MT_ARG_CHECK(device >= 0 && device < ndev, "mt_ctx_create: device %d not in [0,%d)", device, ndev);
    MT_HIP_CHECK(hipSetDevice(device));
    int32_t rc = mt_noalloc_setup(device); // once per device, with the device current
    if (rc) return rc;
    mt_context* c = new (std::nothrow) mt_context();
    if (!c) {
        mt_set_error("mt_ctx_create: out of host memory");
        return MT_ERR_ALLOC;
    }
    c->device = device;
    c->scratch = nullptr;
    c->scratch_bytes = 0;
    c->temp = nullptr;
    c->temp_bytes = 0;
    c->side = nullptr;
    c->nfj = 0;
    if (!(flags & MT_CTX_OWN_STREAM)) {
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            mt_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
            return MT_ERR_HIP;
        }
        c->own_stream = true;
    }
    *out = c;
    return MT_OK;
}

extern "C" int32_t mt_ctx_destroy(mt_context* ctx)
{
    if (!ctx) return MT_OK;
    MtFirstError fe;
    fe(hipSetDevice(ctx->device));
    fe(hipStreamSynchronize(ctx->stream));
    if (ctx->scratch) fe(hipFree(ctx->scratch));
    if (ctx->temp) fe(hipFree(ctx->temp));
    if (ctx->side) {
        fe(hipStreamSynchronize(ctx->side));
        fe(hipStreamDestroy(ctx->side));
    }
    for (int k = 0; k < ctx->nfj; k++) {
        fe(hipEventDestroy(ctx->fork[k]));
        fe(hipEventDestroy(ctx->join[k]));
    }
    if (ctx->own_stream) fe(hipStreamDestroy(ctx->stream));
    delete ctx;
    return fe.status("mt_ctx_destroy");
}

int32_t mt_ctx_side_stream(mt_context* ctx, hipStream_t* out)
{
    if (!ctx->side) {
        hipError_t e = hipSetDevice(ctx->device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
        if (e != hipSuccess) {
            ctx->side = nullptr;
            mt_set_error("side stream: hipStreamCreate failed: %s", hipGetErrorString(e));
            return MT_ERR_HIP;
        }
    }
    for (; ctx->nfj < 64; ctx->nfj++) {
        if (hipEventCreateWithFlags(&ctx->fork[ctx->nfj], hipEventDisableTiming) != hipSuccess) {
            mt_set_error("side stream: hipEventCreate failed");
            return MT_ERR_HIP;
        }
        if (hipEventCreateWithFlags(&ctx->join[ctx->nfj], hipEventDisableTiming) != hipSuccess) {
            (void)hipEventDestroy(ctx->fork[ctx->nfj]);
            mt_set_error("side stream: hipEventCreate failed");
            return MT_ERR_HIP;
        }
    }
    *out = ctx->side;
    return MT_OK;
}

extern "C" int32_t mt_ctx_sync(mt_context* ctx)
{
    MT_ARG_CHECK(ctx != nullptr, "mt_ctx_sync: ctx is NULL");
    MT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return MT_OK;
}

int32_t mt_ctx_temp(mt_context* ctx, size_t bytes, void** out)
{
    if (bytes <= ctx->temp_bytes) {
        *out = ctx->temp;
        return MT_OK;
    }
    size_t want = bytes < (size_t)1 << 20 ? (size_t)1 << 20 : bytes;
    MT_HIP_CHECK(hipStreamSynchronize(ctx->stream)); // earlier work may still read the old buffer
    if (ctx->temp) {
        MT_HIP_CHECK(hipFree(ctx->temp));
        ctx->temp = nullptr;
        ctx->temp_bytes = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        mt_set_error("temporary allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
        return MT_ERR_ALLOC;
    }
    ctx->temp = p;
    ctx->temp_bytes = want;
    *out = p;
    return MT_OK;
}

int32_t mt_ctx_scratch(mt_context* ctx, size_t bytes, void** out)
{
    if (bytes <= ctx->scratch_bytes) {
        *out = ctx->scratch;
        return MT_OK;
    }
    size_t want = bytes < (size_t)1 << 20 ? (size_t)1 << 20 : bytes;
    MT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->scratch) {
        MT_HIP_CHECK(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        mt_set_error("scratch allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
        return MT_ERR_ALLOC;
    }
    // Status words start at epoch 0 and so does the device-side epoch
    // counter; every scan draws a fresh non-zero epoch.  Cleared with a fill
    // kernel rather than a memset node.
    uint64_t words = want / 4;
    int blocks = (int)((words + 255) / 256 < 65536 ? (words + 255) / 256 : 65536);
    k_fill_i32<<<blocks, 256, 0, ctx->stream>>>((int32_t*)p, 0, words);
    MT_LAUNCH_CHECK();
    ctx->scratch = p;
    ctx->scratch_bytes = want;
    *out = p;
    return MT_OK;
}

extern "C" int32_t mt_buf_alloc(mt_context* ctx, uint64_t bytes, void** dptr)
{
    MT_ARG_CHECK(ctx && dptr, "mt_buf_alloc: NULL argument");
    *dptr = nullptr;
    MT_HIP_CHECK(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 4);
    if (e != hipSuccess) {
        mt_set_error("hipMalloc(%llu) failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
        return MT_ERR_ALLOC;
    }
    return MT_OK;
}

extern "C" int32_t mt_buf_free(mt_context* ctx, void* dptr)
{
    MT_ARG_CHECK(ctx != nullptr, "mt_buf_free: ctx is NULL");
    if (!dptr) return MT_OK;
    MT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    MT_HIP_CHECK(hipFree(dptr));
    return MT_OK;
}

extern "C" int32_t mt_buf_upload(mt_context* ctx, void* dst, const void* src, uint64_t bytes)
{
    MT_ARG_CHECK(ctx && dst && src, "mt_buf_upload: NULL argument");
    MT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    MT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return MT_OK;
}

extern "C" int32_t mt_buf_download(mt_context* ctx, void* dst, const void* src, uint64_t bytes)
{
    MT_ARG_CHECK(ctx && dst && src, "mt_buf_download: NULL argument");
    MT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return MT_OK;
}

extern "C" int32_t mt_buf_fill_i32(mt_context* ctx, int32_t* dst, int32_t value, uint64_t count)
{
    MT_ARG_CHECK(ctx && dst, "mt_buf_fill_i32: NULL argument");
    if (count == 0) return MT_OK;
    int blocks = (int)((count + 255) / 256 < 65536 ? (count + 255) / 256 : 65536);
    k_fill_i32<<<blocks, 256, 0, ctx->stream>>>(dst, value, count);
    MT_LAUNCH_CHECK();
    return MT_OK;
}
