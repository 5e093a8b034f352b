// FCT stage kernels for one patch (Rectangle::FCTTimeStep, Rectangle.cpp:1255-1623).
//
// Layout: every plane of a patch is (n_x + 4) rows of `pitch` doubles, element
// (i, j) at (i + 2) * pitch + (j + 2), p contiguous, so a 64-lane wavefront
// reads 64 consecutive p-cells (512 B) per access.
//
// Entries the reference never writes (outside each quantity's loop range)
// are zero in the reference (vectors are zero-assigned in the constructor,
// Rectangle.cpp:10-38); the kernels reproduce that by computing each quantity
// only inside its range and using 0 outside.
#include "vx_dev.h"
#include "vx_internal.h"

using namespace vxi;
using namespace vxd;

namespace {

constexpr int TI = 16;  // x rows per tile
constexpr int TJ = 64;  // p columns per tile (one wavefront)

struct Sub0Args {
    vx_patch p;
    vx_fields fl;
    int step;
    int ns;          // step + 1 terms in the RK combination
    double a[6];     // RK coefficients times dt (Rectangle.cpp:1396-1517)
    double aSum;     // a[0] + ... + a[ns-1], summed left to right
};

__device__ __forceinline__ double load_or0(const double *pl, int pitch, int nx, int np, int i,
                                           int j) {
    return (i >= -2 && i < nx + 2 && j >= -2 && j < np + 2) ? pl[pidx(pitch, i, j)] : 0.0;
}

// EMFieldSolver::GetASquared(GetFinestIndex(i)) scaled by q*q (Rectangle.cpp:1280)
__device__ __forceinline__ double face_as(const Sub0Args &A, int i) {
    long fi = (long)A.p.rtb * (long)(i + A.p.x_pos);
    long c = fi < 0 ? 0 : (fi > A.fl.n_f ? A.fl.n_f : fi);
    return A.p.q * A.p.q * A.fl.a2[c];
}

// Rectangle::GetEfield(i) (Rectangle.cpp:1055-1067); E is periodic on the
// finest grid (EMSolver.cpp:137-154 wraps every stencil index).
__device__ __forceinline__ double patch_E(const vx_patch &p, const vx_fields &fl, int i) {
    return vxd::patch_E(fl.E, fl.n_f, p.rtb, p.x_pos, i);
}

// Sub-step 0: fluxes of stage `step`, RK combination.  One 256-thread block
// covers NS positions i in [i0, i0+16), j in [j0, j0+64).
__global__ __launch_bounds__(256) void k_sub0(Sub0Args A) {
    const vx_patch &P = A.p;
    const int nx = P.n_x, np = P.n_p, pitch = P.pitch;
    const int i0 = -2 + (int)blockIdx.y * TI;
    const int j0 = -2 + (int)blockIdx.x * TJ;
    const int tid = threadIdx.x;

    __shared__ double F1[TI + 3][TJ + 4];  // f slot 1, i in [i0-2, i0+17), j in [j0-2, j0+66)
    __shared__ double G[TI + 3][TJ + 4];   // Gamma(p_j, as_i), i in [i0-1, i0+18), j in [j0-1, j0+67)
    __shared__ double EX[TI][TJ + 2], FX[TI][TJ + 2];          // x-faces, j in [j0-1, j0+65)
    __shared__ double EP[TI + 2][TJ], FP[TI + 2][TJ];          // p-faces, i in [i0-1, i0+17)
    __shared__ double AS[TI + 3];  // q^2 A^2 at x-faces i0-1 .. i0+17
    __shared__ double EM[TI + 2];  // q E at columns i0-1 .. i0+16

    const double *f1 = P.f[1];
    for (int e = tid; e < (TI + 3) * (TJ + 3); e += 256) {
        int r = e / (TJ + 3), c = e % (TJ + 3);
        F1[r][c] = load_or0(f1, pitch, nx, np, i0 - 2 + r, j0 - 2 + c);
    }
    if (tid < TI + 3) AS[tid] = face_as(A, i0 - 1 + tid);
    if (tid >= 64 && tid < 64 + TI + 2) {
        int r = tid - 64, i = i0 - 1 + r;
        EM[r] = (i >= -1 && i < nx + 1) ? P.q * patch_E(P, A.fl, i) : 0.0;  // columns ep reads
    }
    __syncthreads();
    for (int e = tid; e < (TI + 3) * (TJ + 3); e += 256) {
        int r = e / (TJ + 3), c = e % (TJ + 3);
        double pj = momentum(P.pmin, P.dp, j0 - 1 + c, P.p_pos);
        G[r][c] = gamma_(pj, AS[r], P.mc2inv);
    }
    __syncthreads();
    const double dpcc = P.dp_inv * P.cc;
    const double ccdx = P.cc * P.dx_inv;
    // x-face velocities and WENO values (Rectangle.cpp:1278-1293)
    for (int e = tid; e < TI * (TJ + 2); e += 256) {
        int r = e / (TJ + 2), c = e % (TJ + 2);
        int i = i0 + r, j = j0 - 1 + c;
        double ex = 0.0, fx = 0.0;
        if (i >= -1 && i < nx + 2 && j >= -1 && j < np + 1)
            ex = dpcc * (G[r + 1][c + 1] - G[r + 1][c]);
        if (i >= 0 && i < nx + 1 && j >= -1 && j < np + 1)
            fx = weno(F1[r][c + 1], F1[r + 1][c + 1], F1[r + 2][c + 1], F1[r + 3][c + 1], ex > 0.0);
        EX[r][c] = ex;
        FX[r][c] = fx;
    }
    // p-face forces and WENO values (Rectangle.cpp:1294-1312)
    for (int e = tid; e < (TI + 2) * TJ; e += 256) {
        int r = e / TJ, c = e % TJ;
        int i = i0 - 1 + r, j = j0 + c;
        double ep = 0.0, fp = 0.0;
        if (i >= -1 && i < nx + 1 && j >= -1 && j < np + 2)
            ep = EM[r] - ccdx * (G[r + 1][c + 1] - G[r][c + 1]);
        if (i >= -1 && i < nx + 1 && j >= 0 && j < np + 1)
            fp = weno(F1[r + 1][c], F1[r + 1][c + 1], F1[r + 1][c + 2], F1[r + 1][c + 3], ep > 0.0);
        EP[r][c] = ep;
        FP[r][c] = fp;
    }
    __syncthreads();

    const double w3 = 1 / 48.0;
    const int c = tid & 63;
    const int j = j0 + c;
    for (int rr = tid >> 6; rr < TI; rr += 4) {
        const int i = i0 + rr;
        if (i >= nx + 2 || j >= np + 2) continue;
        const long id = pidx(pitch, i, j);
        // high- and low-order x flux (Rectangle.cpp:1313-1352)
        double fxh = 0.0, fxl = 0.0;
        if (i >= 0 && i < nx + 1 && j >= -1 && j < np + 1) {
            double am = EX[rr][c + 1], ap1 = EX[rr][c + 2], am1 = EX[rr][c];
            double fm = FX[rr][c + 1], fp1 = FX[rr][c + 2], fm1 = FX[rr][c];
            fxh = P.dx_inv * (fm * am + w3 * (fp1 - fm1) * (ap1 - am1));
            fxl = P.dx_inv * ((am > 0.0 ? F1[rr + 1][c + 2] : F1[rr + 2][c + 2]) * am);
            P.FxH[A.step][id] = fxh;
            if (A.step == 0) P.FxL0[id] = fxl;
        }
        // high- and low-order p flux (Rectangle.cpp:1353-1394)
        double fph = 0.0, fpl = 0.0;
        if (i >= -1 && i < nx + 1 && j >= 0 && j < np + 1) {
            double am = EP[rr + 1][c], ap1 = EP[rr + 2][c], am1 = EP[rr][c];
            double fm = FP[rr + 1][c], fp1 = FP[rr + 2][c], fm1 = FP[rr][c];
            fph = P.dp_inv * (fm * am + w3 * (fp1 - fm1) * (ap1 - am1));
            fpl = P.dp_inv * ((am > 0.0 ? F1[rr + 2][c + 1] : F1[rr + 2][c + 2]) * am);
            P.FpH[A.step][id] = fph;
            if (A.step == 0) P.FpL0[id] = fpl;
        }
        // RK flux combination (Rectangle.cpp:1396-1517), all NS positions
        double lx = A.step == 0 ? fxl : P.FxL0[id];
        double lp = A.step == 0 ? fpl : P.FpL0[id];
        double xLS = A.aSum * lx;
        double pLS = A.aSum * lp;
        double sx = A.a[0] * (A.step == 0 ? fxh : P.FxH[0][id]);
        double sp = A.a[0] * (A.step == 0 ? fph : P.FpH[0][id]);
        for (int k = 1; k < A.ns; k++) {
            double hx = (k == A.step) ? fxh : P.FxH[k][id];
            double hp = (k == A.step) ? fph : P.FpH[k][id];
            sx = sx + A.a[k] * hx;
            sp = sp + A.a[k] * hp;
        }
        P.FxDS[id] = sx - xLS;
        P.FpDS[id] = sp - pLS;
    }
}

// Low-order predictor f2 on interior cells (Rectangle.cpp:1518-1534) in
// gather form.  For cell (i, j) the serial scatter loop touches it in the
// order: iteration (i,j) adds FxLS(i,j) then FpLS(i,j); iteration (i,j+1)
// subtracts FpLS(i,j+1); iteration (i+1,j) subtracts FxLS(i+1,j) -- each only
// when that iteration lies in the loop ranges [xm,xp) x [pm,pp).
struct LoopRange {
    int xm, xp, pm, pp;
    __device__ bool in(int i, int j) const { return i >= xm && i < xp && j >= pm && j < pp; }
};

__device__ __forceinline__ LoopRange loop_range(const vx_patch &P) {
    LoopRange L;
    L.xm = P.left ? 1 : 0;
    L.xp = P.right ? P.n_x : P.n_x + 1;
    L.pp = P.up ? P.n_p : P.n_p + 1;
    L.pm = P.down ? 1 : 0;
    return L;
}

__global__ __launch_bounds__(256) void k_lo_predictor(vx_patch P, double aSum) {
    int j = blockIdx.x * 64 + (threadIdx.x & 63);
    int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= P.n_x || j >= P.n_p) return;
    LoopRange L = loop_range(P);
    const int pitch = P.pitch;
    double t = P.f[0][pidx(pitch, i, j)];
    if (L.in(i, j)) {
        t += aSum * P.FxL0[pidx(pitch, i, j)];
        t += aSum * P.FpL0[pidx(pitch, i, j)];
    }
    if (L.in(i, j + 1)) t -= aSum * P.FpL0[pidx(pitch, i, j + 1)];
    if (L.in(i + 1, j)) t -= aSum * P.FxL0[pidx(pitch, i + 1, j)];
    P.f[2][pidx(pitch, i, j)] = t;
}

// Zalesak R+- for i in [-1, n_x+1), j in [-1, n_p+1) (Rectangle.cpp:1538-1579)
__global__ __launch_bounds__(256) void k_limiter_r(vx_patch P) {
    int j = -1 + (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    int i = -1 + (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (i >= P.n_x + 1 || j >= P.n_p + 1) return;
    const int pitch = P.pitch;
    const double *FxDS = P.FxDS, *FpDS = P.FpDS, *f0 = P.f[0], *f2 = P.f[2];
    long idx = pidx(pitch, i, j);
    double Pp = valmax(0.0, FxDS[idx]) - valmin(0.0, FxDS[pidx(pitch, i + 1, j)]) +
                valmax(0.0, FpDS[idx]) - valmin(0.0, FpDS[pidx(pitch, i, j + 1)]);
    double Pm = valmax(0.0, FxDS[pidx(pitch, i + 1, j)]) - valmin(0.0, FxDS[idx]) +
                valmax(0.0, FpDS[pidx(pitch, i, j + 1)]) - valmin(0.0, FpDS[idx]);
    double a0 = f0[idx], a2 = f2[idx];
    double b0 = f0[pidx(pitch, i + 1, j)], b2 = f2[pidx(pitch, i + 1, j)];
    double c0 = f0[pidx(pitch, i - 1, j)], c2 = f2[pidx(pitch, i - 1, j)];
    double d0 = f0[pidx(pitch, i, j + 1)], d2 = f2[pidx(pitch, i, j + 1)];
    double e0 = f0[pidx(pitch, i, j - 1)], e2 = f2[pidx(pitch, i, j - 1)];
    double wMax = valmax(valmax(a0, a2), valmax(valmax(b0, b2), valmax(valmax(c0, c2),
                  valmax(valmax(d0, d2), valmax(e0, e2)))));
    double wMin = valmin(valmin(a0, a2), valmin(valmin(b0, b2), valmin(valmin(c0, c2),
                  valmin(valmin(d0, d2), valmin(e0, e2)))));
    double Qm = -wMin + a2;
    double Qp = wMax - a2;
    P.Rp[idx] = Pp > 0.0 ? valmin(1.0, Qp / Pp) : 0.0;
    P.Rm[idx] = Pm > 0.0 ? valmin(1.0, Qm / Pm) : 0.0;
}

// C on the NS range: 1 everywhere, the limiter formula on interior faces
// i in [1, n_x), j in [0, n_p) (Rectangle.cpp:1581-1594)
__global__ __launch_bounds__(256) void k_limiter_c(vx_patch P) {
    int j = -2 + (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    int i = -2 + (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (i >= P.n_x + 2 || j >= P.n_p + 2) return;
    const int pitch = P.pitch;
    long idx = pidx(pitch, i, j);
    double cx = 1.0, cp = 1.0;
    if (i >= 1 && i < P.n_x && j >= 0 && j < P.n_p) {
        long idxi = pidx(pitch, i - 1, j), idxj = pidx(pitch, i, j - 1);
        cx = P.FxDS[idx] > 0.0 ? valmin(P.Rp[idx], P.Rm[idxi]) : valmin(P.Rp[idxi], P.Rm[idx]);
        cp = P.FpDS[idx] > 0.0 ? valmin(P.Rp[idx], P.Rm[idxj]) : valmin(P.Rp[idxj], P.Rm[idx]);
    }
    P.Cx[idx] = cx;
    P.Cp[idx] = cp;
}

// f1 = f2 + limited antidiffusive fluxes on interior cells
// (Rectangle.cpp:1596-1612), same gather order as the predictor.
__global__ __launch_bounds__(256) void k_corrector(vx_patch P) {
    int j = blockIdx.x * 64 + (threadIdx.x & 63);
    int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= P.n_x || j >= P.n_p) return;
    LoopRange L = loop_range(P);
    const int pitch = P.pitch;
    long idx = pidx(pitch, i, j);
    double t = P.f[2][idx];
    if (L.in(i, j)) {
        t += P.Cx[idx] * P.FxDS[idx];
        t += P.Cp[idx] * P.FpDS[idx];
    }
    if (L.in(i, j + 1)) {
        long k = pidx(pitch, i, j + 1);
        t -= P.Cp[k] * P.FpDS[k];
    }
    if (L.in(i + 1, j)) {
        long k = pidx(pitch, i + 1, j);
        t -= P.Cx[k] * P.FxDS[k];
    }
    P.f[1][idx] = t;
}

// zero the 2-wide ghost ring of one slot (sentinel neighbours on all sides)
__global__ void k_ghost_zero(double *g, int nx, int np, int pitch) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int ring_rows = 2 * (np + 4);  // rows i=-2,-1 and n_x, n_x+1 (full width)
    int side = 2 * 2 * nx;         // columns j=-2,-1,n_p,n_p+1 for i in [0,n_x)
    if (t < 2 * ring_rows) {
        int r = t / (np + 4), c = t % (np + 4);
        int i = r < 2 ? r - 2 : nx + (r - 2);
        g[pidx(pitch, i, c - 2)] = 0.0;
    } else if (t < 2 * ring_rows + side) {
        int u = t - 2 * ring_rows;
        int i = u / 4, k = u % 4;
        int j = k < 2 ? k - 2 : np + (k - 2);
        g[pidx(pitch, i, j)] = 0.0;
    }
}

// f0 <- f1 over the NS range (Rectangle.cpp:1614-1622)
__global__ __launch_bounds__(256) void k_copy_plane(const double *src, double *dst, long n) {
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e < n) dst[e] = src[e];
}

void rk_coefficients(int step, double dt, double *a, int &ns, double &aSum) {
    ns = step + 1;
    for (int k = 0; k < 6; k++) a[k] = 0.0;
    for (int k = 0; k < ns; k++) a[k] = vx::rk_a[step][k] * dt;
    aSum = a[0];
    for (int k = 1; k < ns; k++) aSum = aSum + a[k];
}

dim3 grid_cells(int nx, int np, int extra_i, int extra_j) {
    return dim3((unsigned)((np + extra_j + 63) / 64), (unsigned)((nx + extra_i + 3) / 4));
}

}  // namespace

extern "C" {

int vx_fct_sub0_fluxes(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (step < 0 || step > 5 || !fl) return fail(ctx, VX_EINVAL, "sub0: bad step %d", step);
    if ((rc = validate_span(ctx, p, fl->n_f))) return rc;
    Sub0Args A;
    A.p = *p;
    A.fl = *fl;
    A.step = step;
    rk_coefficients(step, dt, A.a, A.ns, A.aSum);
    dim3 grid((unsigned)((p->n_p + 4 + TJ - 1) / TJ), (unsigned)((p->n_x + 4 + TI - 1) / TI));
    {
        LaunchTimer lt(ctx, "k_sub0");
        k_sub0<<<grid, 256, 0, ctx->stream>>>(A);
    }
    return check_launch(ctx, "k_sub0");
}

int vx_fct_sub0_predictor(vx_ctx *ctx, const vx_patch *p, int step, double dt) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (step < 0 || step > 5) return fail(ctx, VX_EINVAL, "sub0 predictor: bad step %d", step);
    double a[6], aSum;
    int ns;
    rk_coefficients(step, dt, a, ns, aSum);
    k_lo_predictor<<<grid_cells(p->n_x, p->n_p, 0, 0), 256, 0, ctx->stream>>>(*p, aSum);
    return check_launch(ctx, "k_lo_predictor");
}

int vx_fct_sub0(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl) {
    int rc = vx_fct_sub0_fluxes(ctx, p, step, dt, fl);
    if (rc) return rc;
    return vx_fct_sub0_predictor(ctx, p, step, dt);
}

int vx_fct_sub1(vx_ctx *ctx, const vx_patch *p) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    k_limiter_r<<<grid_cells(p->n_x, p->n_p, 2, 2), 256, 0, ctx->stream>>>(*p);
    rc = check_launch(ctx, "k_limiter_r");
    if (rc) return rc;
    k_limiter_c<<<grid_cells(p->n_x, p->n_p, 4, 4), 256, 0, ctx->stream>>>(*p);
    return check_launch(ctx, "k_limiter_c");
}

int vx_fct_sub2(vx_ctx *ctx, const vx_patch *p) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    k_corrector<<<grid_cells(p->n_x, p->n_p, 0, 0), 256, 0, ctx->stream>>>(*p);
    return check_launch(ctx, "k_corrector");
}

int vx_fct_sub3(vx_ctx *ctx, const vx_patch *p) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    long n = (long)(p->n_x + 4) * p->pitch;
    k_copy_plane<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p->f[1], p->f[0], n);
    return check_launch(ctx, "k_copy_plane");
}

int vx_ghost_zero(vx_ctx *ctx, const vx_patch *p, int slot) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (slot < 0 || slot > 2) return fail(ctx, VX_EINVAL, "ghost: bad slot %d", slot);
    int total = 4 * (p->n_p + 4) + 4 * p->n_x;
    k_ghost_zero<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(p->f[slot], p->n_x,
                                                                          p->n_p, p->pitch);
    return check_launch(ctx, "k_ghost_zero");
}

int vx_fct_stage_walls(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (!(p->left && p->right && p->down && p->up))
        return fail(ctx, VX_EINVAL, "stage_walls: patch must touch all four domain edges");
    if ((rc = vx_fct_sub0(ctx, p, step, dt, fl))) return rc;
    if ((rc = vx_ghost_zero(ctx, p, 2))) return rc;
    if ((rc = vx_fct_sub1(ctx, p))) return rc;
    if ((rc = vx_fct_sub2(ctx, p))) return rc;
    if ((rc = vx_ghost_zero(ctx, p, 1))) return rc;
    if (step == 5) rc = vx_fct_sub3(ctx, p);
    return rc;
}

}  // extern "C"
