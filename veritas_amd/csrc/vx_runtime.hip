// Context, memory and timing entry points of the C ABI.
#include "vx_internal.h"

using namespace vxi;

extern "C" {

int vx_ctx_create(int device, vx_ctx **out) {
    if (!out) return VX_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return VX_EHIP;
    vx_ctx *c = new vx_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return VX_EHIP;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return VX_EHIP;
        }
    *out = c;
    return VX_OK;
}

void vx_ctx_destroy(vx_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->tev) (void)hipEventDestroy(e);
    if (ctx->mom_table) (void)hipFree(ctx->mom_table);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *vx_last_error(vx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int vx_sync(vx_ctx *ctx) { return check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync"); }

void *vx_stream(vx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int vx_malloc(vx_ctx *ctx, size_t bytes, void **dptr) {
    if (!dptr) return fail(ctx, VX_EINVAL, "null out pointer");
    *dptr = nullptr;
    if (bytes == 0) return VX_OK;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) return fail(ctx, VX_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipMemsetAsync(*dptr, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return check_hip(ctx, e, "zero fill");
}

int vx_free(vx_ctx *ctx, void *dptr) {
    if (!dptr) return VX_OK;
    (void)hipStreamSynchronize(ctx->stream);
    return check_hip(ctx, hipFree(dptr), "hipFree");
}

int vx_memcpy_h2d(vx_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "h2d");
    return rc ? rc : vx_sync(ctx);
}

int vx_memcpy_d2h(vx_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    return rc ? rc : vx_sync(ctx);
}

int vx_memcpy_d2d(vx_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream), "d2d");
}

__global__ void k_fill(double *dst, double v, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

int vx_fill(vx_ctx *ctx, double *dst, double value, size_t count) {
    if (count == 0) return VX_OK;
    size_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_fill<<<(unsigned)blocks, 256, 0, ctx->stream>>>(dst, value, count);
    return check_launch(ctx, "fill");
}

int vx_event_record(vx_ctx *ctx, int slot) {
    if (slot < 0 || slot >= 8) return fail(ctx, VX_EINVAL, "event slot %d", slot);
    return check_hip(ctx, hipEventRecord(ctx->ev[slot], ctx->stream), "event record");
}

int vx_event_elapsed_ms(vx_ctx *ctx, int a, int b, float *ms) {
    if (a < 0 || a >= 8 || b < 0 || b >= 8 || !ms) return fail(ctx, VX_EINVAL, "event slots");
    int rc = check_hip(ctx, hipEventSynchronize(ctx->ev[b]), "event sync");
    if (rc) return rc;
    return check_hip(ctx, hipEventElapsedTime(ms, ctx->ev[a], ctx->ev[b]), "event elapsed");
}

int vx_timing_begin(vx_ctx *ctx, const char *kernel) {
    ctx->timed_kernel = kernel ? kernel : "";
    ctx->tcount = 0;
    return VX_OK;
}

int vx_timing_end(vx_ctx *ctx, double *total_ms, int *launches) {
    int rc = vx_sync(ctx);
    if (rc) return rc;
    double tot = 0.0;
    for (size_t k = 0; k < ctx->tcount; k++) {
        float ms = 0.0f;
        rc = check_hip(ctx, hipEventElapsedTime(&ms, ctx->tev[2 * k], ctx->tev[2 * k + 1]), "timing");
        if (rc) return rc;
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (int)ctx->tcount;
    ctx->timed_kernel.clear();
    ctx->tcount = 0;
    return VX_OK;
}

int32_t vx_plane_pitch(int32_t n_p) { return ((n_p + 4 + 7) / 8) * 8; }

size_t vx_plane_elems(int32_t n_x, int32_t n_p) {
    return (size_t)(n_x + 4) * (size_t)vx_plane_pitch(n_p);
}

}  // extern "C"
