// Shared internals of the HIP library: context, error reporting, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <array>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/veritas_amd.h"

struct vx_ctx {
    int device = 0;
    int n_cu = 256;  // compute units of the device (launch-shape choices)
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
    // per-launch timing of one named kernel (vx_timing_begin / vx_timing_end)
    std::string timed_kernel;
    std::vector<hipEvent_t> tev;  // pairs: start, stop
    size_t tcount = 0;
    // p-moment logarithm table of the patch being reduced (vx_moments_patch)
    double *mom_table = nullptr;
    size_t mom_table_n = 0;
    // strip heights of the fused stage chosen per patch shape and stage:
    // {n_x, n_p, rows, mixed, stage}
    std::vector<std::array<int, 5>> ti_cache;
    // device bytes held through vx_malloc (vx_ctx_live_bytes)
    std::unordered_map<void *, size_t> allocs;
    size_t live_bytes = 0;
    // cross-context ordering (vx_ctx_wait) and the asynchronous CFL bound
    hipEvent_t xev = nullptr, cfl_ev = nullptr;
    double *cfl_host = nullptr;  // pinned
    double *cfl_dev = nullptr;
    double cfl_c_dx = 0.0;       // c / dx_f of the pending bound
    bool cfl_pending = false;
    // staging of batched stage launches (vx_stage.hip), freed by its own function
    void *batch_ring = nullptr;
    void (*batch_ring_free)(void *) = nullptr;
    // step-graph recording (vx_graph_begin / vx_graph_end)
    bool capturing = false;
    // this device's fixed work array of vx_poisson_noalloc (vx_field.hip), from
    // hipGetSymbolAddress at context creation; null if that failed
    double *poisson_scratch = nullptr;
    // pinned staging ring of vx_upload_small
    static constexpr int RING = 64, RING_BYTES = 256;
    char *ring = nullptr;
    hipEvent_t ring_ev[RING] = {};
    bool ring_pending[RING] = {};
    int ring_next = 0;
};

// vx_field.hip: looks up this device's fixed work array (ctx->device current)
int vx_noalloc_scratch_init(vx_ctx *ctx);

struct vx_graph {
    hipGraphExec_t exec = nullptr;
};

// Device table of a mesh's patches (vx_amr.hip), also the work list of the
// batched sub-step launches (vx_fct.hip): batch_d holds, per patch, where its
// workgroups start in each batched launch and its strip shape.
struct vx_batch_info {
    int start[5];        // first workgroup of the patch: sub-step-0 tiles, cells +0, limiter tiles, cells +4, plane copy
    int TI, nstrip, nbx; // strips of the streaming sub-step-0 kernel
    int lTI, lnstrip, lnbx;  // strips of the streaming limiter
};
struct vx_patch_table {
    vx_ctx *ctx = nullptr;
    vx_patch *d = nullptr;
    int32_t n = 0;
    uint64_t serial = 0;          // identity for the program bounds check
    std::vector<vx_patch> host;   // shapes, for the bounds check
    vx_batch_info *batch_d = nullptr;  // n + 1 entries (the last one: totals)
    std::vector<vx_batch_info> batch;  // host copy
    // strip-height policies of the batched sub-step-0 flux and limiter
    // launches (vx_fct.hip): a positive value is one strip height for every
    // patch, a negative one -W the tallest strips that still give W wavefronts
    // per patch; chosen by timing on the table's first launch of each
    int sub0_pol = -2048, lim_pol = -4096;
    bool sub0_tuned[2] = {false, false}, lim_tuned[2] = {false, false};  // per precision
};

struct vx_moment_plan {
    int n_species = 0, n_levels = 0, n_f = 0;
    void *tables = nullptr;  // device: MomentSrc[] then level_start[]
    size_t src_off = 0, ls_off = 0;
};

namespace vxi {

inline int fail(vx_ctx *ctx, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

inline int check_hip(vx_ctx *ctx, hipError_t e, const char *what) {
    if (e != hipSuccess) return fail(ctx, VX_EHIP, "%s: %s", what, hipGetErrorString(e));
    return VX_OK;
}

inline int check_launch(vx_ctx *ctx, const char *what) {
    return check_hip(ctx, hipGetLastError(), what);
}

// Events around one launch when `name` is the kernel being timed.
struct LaunchTimer {
    vx_ctx *ctx;
    hipEvent_t stop = nullptr;
    LaunchTimer(vx_ctx *c, const char *name) : ctx(c) {
        if (ctx->capturing || ctx->timed_kernel.empty() || ctx->timed_kernel != name) return;
        size_t k = 2 * ctx->tcount;
        if (k + 2 > ctx->tev.size()) {
            for (int e = 0; e < 64; e++) {
                hipEvent_t ev;
                if (hipEventCreate(&ev) != hipSuccess) return;
                ctx->tev.push_back(ev);
            }
        }
        if (hipEventRecord(ctx->tev[k], ctx->stream) != hipSuccess) return;
        stop = ctx->tev[k + 1];
        ctx->tcount++;
    }
    ~LaunchTimer() {
        if (stop) (void)hipEventRecord(stop, ctx->stream);
    }
};

// Plane index for element (i, j) of a patch plane.
__host__ __device__ inline long pidx(int pitch, int i, int j) {
    return (long)(i + 2) * pitch + (j + 2);
}

// Whether iteration (i, j) of the reference's FCT update loops, which run over
// [xm, xp) x [pm, pp) with xm = 1 at a left wall (else 0), xp = n_x at a right
// wall (else n_x + 1), and likewise in p, touches the patch (Rectangle.cpp:
// 1518-1534, 1596-1612).
__host__ __device__ inline bool in_update_loop(const vx_patch &P, int i, int j) {
    return i >= (P.left ? 1 : 0) && i < (P.right ? P.n_x : P.n_x + 1) && j >= (P.down ? 1 : 0) &&
           j < (P.up ? P.n_p : P.n_p + 1);
}

// Corrected value of interior cell (i, j) (Rectangle.cpp:1596-1612) in gather
// form: f^td plus the limited antidiffusive fluxes in the order the serial
// scatter loop applies them -- iteration (i, j) adds Cx FxDS and Cp FpDS at
// (i, j), iteration (i, j+1) subtracts Cp FpDS at (i, j+1), iteration (i+1, j)
// subtracts Cx FxDS at (i+1, j), each only inside the loop ranges.
__device__ inline double corrected_cell(const vx_patch &P, int i, int j) {
    const int pitch = P.pitch;
    const long idx = pidx(pitch, i, j);
    double t = P.f[2][idx];
    if (in_update_loop(P, i, j)) {
        t += P.Cx[idx] * P.FxDS[idx];
        t += P.Cp[idx] * P.FpDS[idx];
    }
    if (in_update_loop(P, i, j + 1)) {
        const long k = pidx(pitch, i, j + 1);
        t -= P.Cp[k] * P.FpDS[k];
    }
    if (in_update_loop(P, i + 1, j)) {
        const long k = pidx(pitch, i + 1, j);
        t -= P.Cx[k] * P.FxDS[k];
    }
    return t;
}

inline int validate_patch(vx_ctx *ctx, const vx_patch *p) {
    if (!p) return fail(ctx, VX_EINVAL, "null patch");
    if (p->n_x < 1 || p->n_p < 1 || p->pitch < p->n_p + 4 || (p->pitch % 8) != 0)
        return fail(ctx, VX_EINVAL, "bad patch shape n_x=%d n_p=%d pitch=%d", p->n_x, p->n_p,
                    p->pitch);
    if (p->rtb < 1) return fail(ctx, VX_EINVAL, "bad relativeToBottom %d", p->rtb);
    return VX_OK;
}

// A patch lies inside the finest grid of n_f cells (Mesh::updateHierarchy only
// builds such patches); the stage kernels wrap finest indices once on that basis.
inline int validate_span(vx_ctx *ctx, const vx_patch *p, int n_f) {
    if (n_f < 1 || p->x_pos < 0 || (long)p->rtb * ((long)p->x_pos + (long)p->n_x) > (long)n_f)
        return fail(ctx, VX_EINVAL, "patch x range [%d, %d) x %d outside the finest grid of %d cells",
                    p->x_pos, p->x_pos + p->n_x, p->rtb, n_f);
    return VX_OK;
}

}  // namespace vxi
