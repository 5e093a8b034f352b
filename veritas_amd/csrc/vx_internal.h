// Shared internals of the HIP library: context, error reporting, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/veritas_amd.h"

struct vx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
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

// Plane index for element (i, j) of a patch plane.
__host__ __device__ inline long pidx(int pitch, int i, int j) {
    return (long)(i + 2) * pitch + (j + 2);
}

inline int validate_patch(vx_ctx *ctx, const vx_patch *p) {
    if (!p) return fail(ctx, VX_EINVAL, "null patch");
    if (p->n_x < 1 || p->n_p < 1 || p->pitch < p->n_p + 4 || (p->pitch % 8) != 0)
        return fail(ctx, VX_EINVAL, "bad patch shape n_x=%d n_p=%d pitch=%d", p->n_x, p->n_p,
                    p->pitch);
    if (p->rtb < 1) return fail(ctx, VX_EINVAL, "bad relativeToBottom %d", p->rtb);
    return VX_OK;
}

}  // namespace vxi
