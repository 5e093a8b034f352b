// Context, memory and timing entry points of the C ABI.
#include <cstring>

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
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        c->n_cu = ncu;
    // the fixed work array of vx_poisson_noalloc; on failure that entry point
    // reports the error and every other one works
    (void)vx_noalloc_scratch_init(c);
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
    if (ctx->batch_ring && ctx->batch_ring_free) ctx->batch_ring_free(ctx->batch_ring);
    if (ctx->xev) (void)hipEventDestroy(ctx->xev);
    if (ctx->cfl_ev) (void)hipEventDestroy(ctx->cfl_ev);
    if (ctx->cfl_host) (void)hipHostFree(ctx->cfl_host);
    if (ctx->cfl_dev) (void)hipFree(ctx->cfl_dev);
    for (auto &e : ctx->ring_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ring) (void)hipHostFree(ctx->ring);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *vx_last_error(vx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// entry points that wait for the device or allocate cannot be recorded in a step graph
static int not_recording(vx_ctx *ctx, const char *what) {
    return ctx->capturing ? fail(ctx, VX_EINVAL, "%s while a step graph is being recorded", what) : VX_OK;
}

int vx_sync(vx_ctx *ctx) {
    if (int rc = not_recording(ctx, "sync")) return rc;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
}

void *vx_stream(vx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int vx_malloc(vx_ctx *ctx, size_t bytes, void **dptr) {
    if (!dptr) return fail(ctx, VX_EINVAL, "null out pointer");
    *dptr = nullptr;
    if (int rc = not_recording(ctx, "vx_malloc")) return rc;
    if (bytes == 0) return VX_OK;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) return fail(ctx, VX_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    ctx->allocs[*dptr] = bytes;
    ctx->live_bytes += bytes;
    e = hipMemsetAsync(*dptr, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return check_hip(ctx, e, "zero fill");
}

int vx_free(vx_ctx *ctx, void *dptr) {
    if (!dptr) return VX_OK;
    if (int rc = not_recording(ctx, "vx_free")) return rc;
    auto it = ctx->allocs.find(dptr);
    if (it != ctx->allocs.end()) {
        ctx->live_bytes -= it->second;
        ctx->allocs.erase(it);
    }
    (void)hipStreamSynchronize(ctx->stream);
    return check_hip(ctx, hipFree(dptr), "hipFree");
}

int vx_memcpy_h2d(vx_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (int rc = not_recording(ctx, "vx_memcpy_h2d")) return rc;
    int rc = check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "h2d");
    return rc ? rc : vx_sync(ctx);
}

int vx_memcpy_d2h(vx_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (int rc = not_recording(ctx, "vx_memcpy_d2h")) return rc;
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


int vx_graph_begin(vx_ctx *ctx) {
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "a step graph is already being recorded");
    int rc = check_hip(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                       "begin capture");
    if (rc == VX_OK) ctx->capturing = true;
    return rc;
}

int vx_graph_end(vx_ctx *ctx, vx_graph **out) {
    if (out) *out = nullptr;
    if (!ctx->capturing) return fail(ctx, VX_EINVAL, "no step graph is being recorded");
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (e != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return fail(ctx, VX_EHIP, "end capture: %s", hipGetErrorString(e));
    }
    if (!out) {  // discard
        (void)hipGraphDestroy(graph);
        return VX_OK;
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(ctx, VX_EHIP, "graph instantiate: %s", hipGetErrorString(e));
    vx_graph *g = new vx_graph;
    g->exec = exec;
    *out = g;
    return VX_OK;
}

int vx_graph_launch(vx_ctx *ctx, const vx_graph *g) {
    if (!g || !g->exec) return fail(ctx, VX_EINVAL, "null graph");
    if (int rc = not_recording(ctx, "vx_graph_launch")) return rc;
    return check_hip(ctx, hipGraphLaunch(g->exec, ctx->stream), "graph launch");
}

void vx_graph_destroy(vx_graph *g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    delete g;
}

int vx_upload_small(vx_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes > (size_t)vx_ctx::RING_BYTES) return fail(ctx, VX_EINVAL, "upload of %zu bytes", bytes);
    if (int rc = not_recording(ctx, "vx_upload_small")) return rc;
    if (!ctx->ring) {
        int rc = check_hip(ctx, hipHostMalloc((void **)&ctx->ring, (size_t)vx_ctx::RING * vx_ctx::RING_BYTES),
                           "pinned ring");
        if (rc) return rc;
        for (auto &ev : ctx->ring_ev)
            if ((rc = check_hip(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming), "ring event")))
                return rc;
    }
    const int k = ctx->ring_next;
    ctx->ring_next = (k + 1) % vx_ctx::RING;
    if (ctx->ring_pending[k]) {
        int rc = check_hip(ctx, hipEventSynchronize(ctx->ring_ev[k]), "ring slot wait");
        if (rc) return rc;
    }
    char *slot = ctx->ring + (size_t)k * vx_ctx::RING_BYTES;
    std::memcpy(slot, src, bytes);
    int rc = check_hip(ctx, hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, ctx->stream), "upload");
    if (rc) return rc;
    rc = check_hip(ctx, hipEventRecord(ctx->ring_ev[k], ctx->stream), "ring event record");
    ctx->ring_pending[k] = rc == VX_OK;
    return rc;
}


int vx_ctx_wait(vx_ctx *waiter, vx_ctx *signaller) {
    if (!waiter || !signaller) return VX_EINVAL;
    if (waiter->device != signaller->device) return fail(waiter, VX_EINVAL, "ctx wait across devices");
    if (!signaller->xev) {
        int rc = check_hip(waiter, hipEventCreateWithFlags(&signaller->xev, hipEventDisableTiming), "ctx event");
        if (rc) return rc;
    }
    int rc = check_hip(waiter, hipEventRecord(signaller->xev, signaller->stream), "ctx event record");
    if (rc) return rc;
    return check_hip(waiter, hipStreamWaitEvent(waiter->stream, signaller->xev, 0), "ctx stream wait");
}

size_t vx_ctx_live_bytes(vx_ctx *ctx) { return ctx ? ctx->live_bytes : 0; }


// Slot `slot` of a patch <-> a host array of (n_x + 2g) x (n_p + 2g) doubles,
// row-major (g = 0, 1, 2 ghost layers kept), one strided copy each way.
static int patch_copy(vx_ctx *ctx, const vx_patch *p, int slot, int ghosts, double *host, bool up) {
    if (!p || !host || slot < 0 || slot > 2 || ghosts < 0 || ghosts > 2 || !p->f[slot])
        return fail(ctx, VX_EINVAL, "patch copy: bad slot %d or ghosts %d", slot, ghosts);
    if (int rc = not_recording(ctx, "patch copy")) return rc;
    const size_t w = (size_t)(p->n_p + 2 * ghosts), h = (size_t)(p->n_x + 2 * ghosts);
    double *dev = p->f[slot] + (size_t)(2 - ghosts) * p->pitch + (2 - ghosts);
    hipError_t e = up ? hipMemcpy2DAsync(dev, sizeof(double) * p->pitch, host, sizeof(double) * w,
                                         sizeof(double) * w, h, hipMemcpyHostToDevice, ctx->stream)
                      : hipMemcpy2DAsync(host, sizeof(double) * w, dev, sizeof(double) * p->pitch,
                                         sizeof(double) * w, h, hipMemcpyDeviceToHost, ctx->stream);
    int rc = check_hip(ctx, e, up ? "patch upload" : "patch download");
    return rc ? rc : vx_sync(ctx);
}

int vx_patch_upload(vx_ctx *ctx, const vx_patch *p, int slot, int ghosts, const double *host) {
    return patch_copy(ctx, p, slot, ghosts, const_cast<double *>(host), true);
}

int vx_patch_download(vx_ctx *ctx, const vx_patch *p, int slot, int ghosts, double *host) {
    return patch_copy(ctx, p, slot, ghosts, host, false);
}

}  // extern "C"
