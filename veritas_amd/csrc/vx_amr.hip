// Adaptive-mesh operations as cell-task programs (ghost fill, restriction,
// regrid transfer, limiter-C synchronisation, coarse-fine fluxes) and the
// refinement-flag kernel.
//
// The host records, once per hierarchy, every cell or face operation that the
// reference's per-rectangle loops perform (Rectangle.cpp:314-614, 892-941,
// 1032-1253, 1625-1960) as a vx_cell_task, in the reference's order, and cuts
// the list into batches whose tasks touch disjoint data; one thread runs one
// task, one launch runs one batch.  These are O(patch perimeter) operations;
// the per-cell arithmetic keeps the reference's expressions so that the
// results match the CPU oracle bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vx_dev.h"
#include "vx_internal.h"

using namespace vxi;
using namespace vxd;


// The RESTRICT tasks of a large batch that tile a rectangle of coarse cells
// (dst cells [i0, i0 + ni) x [j0, j0 + nj), coarse coordinates (di + oi,
// dj + oj)) run as that rectangle (k_restrict_rects): one thread per cell, as
// the tasks would, without reading a task per cell.
struct vx_restrict_rect {
    int32_t dst, src, i0, j0, ni, nj, oi, oj;
    int32_t blk0;  // first workgroup of the rectangle in its batch's launch
    int32_t pad;
};

struct vx_program {
    vx_ctx *ctx = nullptr;
    vx_cell_task *d = nullptr;     // the tasks that run as cell tasks
    int32_t *d_batch = nullptr;    // their batch starts on the device (k_program_seq)
    std::vector<int32_t> batch;    // host copy of the batch starts of all tasks
    std::vector<int32_t> cbatch;   // batch starts in d
    std::vector<int32_t> rbatch;   // first rectangle of each batch (n_batches + 1)
    std::vector<int32_t> rblocks;  // workgroups of each batch's rectangle launch
    vx_restrict_rect *d_rect = nullptr;
    std::vector<vx_cell_task> host;
    uint64_t checked_serial = 0;  // table this program was bounds-checked against
    int checked_r = 0;
};

static uint64_t g_table_serial = 0;

namespace {

constexpr int MAXR = 4;

// Rectangle.cpp:76-89
__constant__ double kInterpA[12] = {
    0.104166666666667, -0.708333333333334, 0.708333333333334, -0.104166666666667,
    0.117647058823529, 0.029411764705882,  0.029411764705882, 0.117647058823529,
    -0.083333333333333, 0.166666666666667, -0.166666666666667, 0.083333333333334};

struct ProgArgs {
    const vx_patch *pt;
    const vx_cell_task *task;
    int begin, end;
    int mode, slot, r;
    double inv_r2;           // 1.0 / std::pow(r, 2)      (Rectangle.cpp:389)
    double r2;               // std::pow((double)r, 2.0)  (Rectangle.cpp:325)
    double inv_rr;           // 1.0 / (r * r)             (Rectangle.hpp:141)
    double ref[3 * MAXR];    // interpolatCoefsREF (Rectangle.cpp:91-101)
    // flux programs
    int step, ns;
    double a[6], aSum;
    double b5[6];  // stage 5's coefficients times dt (stage 4 writes its partial sum)
    const double *a2;
    int n_f;
};

__device__ __forceinline__ double &at(double *pl, int pitch, int i, int j) { return pl[pidx(pitch, i, j)]; }
__device__ __forceinline__ double get(const double *pl, int pitch, int i, int j) { return pl[pidx(pitch, i, j)]; }

// GetInterpolantsREF (Rectangle.cpp:121-137), element k of r.  The device
// code of the task programs is compiled per refinement ratio R (2, 3 or 4), so
// the per-element arrays stay in registers.
template <int R>
__device__ __forceinline__ void interp_ref(const ProgArgs &A, double f1, double f2, double f3, double f4,
                                           double f5, double *out) {
    f5 -= f3; f4 -= f3; f2 -= f3; f1 -= f3;
    double a1 = kInterpA[0] * f1 + kInterpA[1] * f2 + kInterpA[2] * f4 + kInterpA[3] * f5;
    double a2 = kInterpA[4] * f1 + kInterpA[5] * f2 + kInterpA[6] * f4 + kInterpA[7] * f5;
    double a3 = kInterpA[8] * f1 + kInterpA[9] * f2 + kInterpA[10] * f4 + kInterpA[11] * f5;
    for (int i = 0; i < R; i++)
        out[i] = A.ref[3 * i] * a1 + A.ref[3 * i + 1] * a2 + A.ref[3 * i + 2] * a3 + f3;
}

// GetWenoValueFromCoarseLevel (Rectangle.cpp:343-415): element e = k * r + l
// of the corrected r x r interpolants of the coarse cell holding (i, j).
template <int R>
__device__ double coarse_element(const ProgArgs &A, const vx_patch &C, const double *g, int i, int j, int e) {
    const int r = R, pitch = C.pitch;
    const int ic = i / r - C.x_pos, jc = j / r - C.p_pos;
    double temps[5][R];
    for (int k = -2; k < 3; k++)
        interp_ref<R>(A, get(g, pitch, ic - 2, jc + k), get(g, pitch, ic - 1, jc + k), get(g, pitch, ic, jc + k),
                   get(g, pitch, ic + 1, jc + k), get(g, pitch, ic + 2, jc + k), temps[k + 2]);
    double sum(0.0), want = 0.0;
    for (int k = 0; k < r; k++) {
        double part[R];
        interp_ref<R>(A, temps[0][k], temps[1][k], temps[2][k], temps[3][k], temps[4][k], part);
        for (int l = 0; l < r; l++) {
            if (k * r + l == e) want = part[l];
            sum += part[l];
        }
    }
    double correction = get(g, pitch, ic, jc) - A.inv_r2 * sum;
    return want + correction;
}

// GetValueFromFinerLevel (Rectangle.cpp:314-327)
template <int R>
__device__ double finer_mean(const ProgArgs &A, const vx_patch &F, const double *g, int i, int j) {
    const int r = R;
    const int i0 = i * r - F.x_pos, j0 = j * r - F.p_pos;
    double temp(0.0);
    for (int k = 0; k < r; k++)
        for (int l = 0; l < r; l++) temp += get(g, F.pitch, i0 + k, j0 + l);
    temp /= A.r2;
    return temp;
}

// std::min of the limiter-C sync (Rectangle.cpp:1646-1905)
__device__ __forceinline__ double mn(double a, double b) { return stdmin(a, b); }

// one limiter-C task (Rectangle.cpp:1625-1706, 1794-1830, 1879-1915)
template <int R>
__device__ void c_task(const ProgArgs &A, const vx_cell_task &T) {
    const vx_patch &N = A.pt[T.dst];  // neighbour (same level) or coarse patch
    const vx_patch &Q = A.pt[T.src];  // caller
    const int np_ = N.pitch, qp = Q.pitch;
    const int di = T.di, dj = T.dj, si = T.si, sj = T.sj, t = T.aux;
    if (T.op == VX_OP_C_SAME) {
        double c;
        if (t == 0)
            c = get(N.FxDS, np_, di, dj) > 0.0 ? mn(get(Q.Rp, qp, si, sj), get(N.Rm, np_, di - 1, dj))
                                               : mn(get(N.Rp, np_, di - 1, dj), get(Q.Rm, qp, si, sj));
        else if (t == 1)
            c = get(N.FxDS, np_, di, dj) > 0.0 ? mn(get(Q.Rm, qp, si - 1, sj), get(N.Rp, np_, di, dj))
                                               : mn(get(N.Rm, np_, di, dj), get(Q.Rp, qp, si - 1, sj));
        else if (t == 2)
            c = get(N.FpDS, np_, di, dj) > 0.0 ? mn(get(Q.Rp, qp, si, sj), get(N.Rm, np_, di, dj - 1))
                                               : mn(get(N.Rp, np_, di, dj - 1), get(Q.Rm, qp, si, sj));
        else
            c = get(N.FpDS, np_, di, dj) > 0.0 ? mn(get(Q.Rm, qp, si, sj - 1), get(N.Rp, np_, di, dj))
                                               : mn(get(N.Rm, np_, di, dj), get(Q.Rp, qp, si, sj - 1));
        double *Nc = t < 2 ? N.Cx : N.Cp, *Qc = t < 2 ? Q.Cx : Q.Cp;
        at(Nc, np_, di, dj) = c;
        at(Qc, qp, si, sj) = c;
    } else if (T.op == VX_OP_C_UPDATE) {
        double *Nc = t < 2 ? N.Cx : N.Cp, *Qc = t < 2 ? Q.Cx : Q.Cp;
        double c = mn(get(Nc, np_, di, dj), get(Qc, qp, si, sj));
        at(Nc, np_, di, dj) = c;
        at(Qc, qp, si, sj) = c;
    } else {  // VX_OP_C_DIFF: N is the coarse patch, Q the finer caller
        const int r = R;
        if (t < 2) {
            double c = get(N.Cx, np_, di, dj);
            for (int k = 0; k < r; k++) {
                bool pos = get(Q.FxDS, qp, si, sj + k) > 0.0;
                double v = t == 0 ? (pos ? get(Q.Rp, qp, si, sj + k) : get(Q.Rm, qp, si, sj + k))
                                  : (pos ? get(Q.Rm, qp, si - 1, sj + k) : get(Q.Rp, qp, si - 1, sj + k));
                c = mn(c, v);
            }
            bool posc = get(N.FxDS, np_, di, dj) > 0.0;
            double vc = t == 0 ? (posc ? get(N.Rm, np_, di - 1, dj) : get(N.Rp, np_, di - 1, dj))
                               : (posc ? get(N.Rp, np_, di, dj) : get(N.Rm, np_, di, dj));
            c = mn(c, vc);
            for (int k = 0; k < r; k++) at(Q.Cx, qp, si, sj + k) = c;
            at(N.Cx, np_, di, dj) = c;
        } else {
            double c = get(N.Cp, np_, di, dj);
            for (int k = 0; k < r; k++) {
                bool pos = get(Q.FpDS, qp, si + k, sj) > 0.0;
                double v = t == 2 ? (pos ? get(Q.Rp, qp, si + k, sj) : get(Q.Rm, qp, si + k, sj))
                                  : (pos ? get(Q.Rm, qp, si + k, sj - 1) : get(Q.Rp, qp, si + k, sj - 1));
                c = mn(c, v);
            }
            bool posc = get(N.FpDS, np_, di, dj) > 0.0;
            double vc = t == 2 ? (posc ? get(N.Rm, np_, di, dj - 1) : get(N.Rp, np_, di, dj - 1))
                               : (posc ? get(N.Rp, np_, di, dj) : get(N.Rm, np_, di, dj));
            c = mn(c, vc);
            for (int k = 0; k < r; k++) at(Q.Cp, qp, si + k, sj) = c;
            at(N.Cp, np_, di, dj) = c;
        }
    }
}

// EMFieldSolver::GetASquared (EMSolver.cpp:133-135) at a patch's x-face
__device__ __forceinline__ double face_a2(const ProgArgs &A, const vx_patch &P, int i) {
    int fi = P.rtb * (i + P.x_pos);
    int c = fi < 0 ? 0 : (fi > A.n_f ? A.n_f : fi);
    return A.a2[c];
}

// RGKGetFluxXL / RGKGetFluxPL of the finer patch (Rectangle.cpp:1216-1253),
// with their own association: dx_inv * f * am
__device__ double lo_flux_x(const ProgArgs &A, const vx_patch &F, int i, int j) {
    double as = F.q * F.q * face_a2(A, F, i);
    double am = F.dp_inv * F.cc *
                (gamma_(momentum(F.pmin, F.dp, j + 1, F.p_pos), as, F.mc2inv) -
                 gamma_(momentum(F.pmin, F.dp, j, F.p_pos), as, F.mc2inv));
    return F.dx_inv * (am > 0.0 ? get(F.f[1], F.pitch, i - 1, j) : get(F.f[1], F.pitch, i, j)) * am;
}
__device__ double lo_flux_p(const ProgArgs &A, const vx_patch &F, const double *E, int i, int j) {
    double as_1 = F.q * F.q * face_a2(A, F, i);
    double as_2 = F.q * F.q * face_a2(A, F, i + 1);
    double Em = F.q * patch_E(E, A.n_f, F.rtb, F.x_pos, i);
    double mom = momentum(F.pmin, F.dp, j, F.p_pos);
    double am = Em - F.cc * F.dx_inv * (gamma_(mom, as_2, F.mc2inv) - gamma_(mom, as_1, F.mc2inv));
    return F.dp_inv * (am > 0.0 ? get(F.f[1], F.pitch, i, j - 1) : get(F.f[1], F.pitch, i, j)) * am;
}

// coarse-fine flux task (Rectangle.cpp:1313-1394 else-branches, then the RK
// combination of Rectangle.cpp:1396-1517 at that face)
template <int R>
__device__ void flux_task(const ProgArgs &A, const double *E, const vx_cell_task &T) {
    const vx_patch &K = A.pt[T.dst];
    const vx_patch &F = A.pt[T.src];
    const int r = R, kp = K.pitch, fp = F.pitch;
    const bool xdir = T.op == VX_OP_FLUX_X;
    double *const *KH = xdir ? K.FxH : K.FpH;
    double *const *FH = xdir ? F.FxH : F.FpH;
    double *KL = xdir ? K.FxL0 : K.FpL0;
    const double *FL = xdir ? F.FxL0 : F.FpL0;
    double h = 0.0;
    for (int k = 0; k < r; k++) h += xdir ? get(FH[A.step], fp, T.si, T.sj + k) : get(FH[A.step], fp, T.si + k, T.sj);
    h *= A.inv_rr;
    at(KH[A.step], kp, T.di, T.dj) = h;
    if (A.step == 0) {
        double l = 0.0;
        for (int k = 0; k < r; k++) {
            int i = xdir ? T.si : T.si + k, j = xdir ? T.sj + k : T.sj;
            double v;
            if ((T.aux >> k) & 1) v = get(FL, fp, i, j);
            else v = xdir ? lo_flux_x(A, F, i, j) : lo_flux_p(A, F, E, i, j);
            l += v;
        }
        l *= A.inv_rr;
        at(KL, kp, T.di, T.dj) = l;
    }
    double s = A.a[0] * get(KH[0], kp, T.di, T.dj);
    for (int k = 1; k < A.ns; k++) s = s + A.a[k] * get(KH[k], kp, T.di, T.dj);
    double ls = A.aSum * get(KL, kp, T.di, T.dj);
    at(xdir ? K.FxDS : K.FpDS, kp, T.di, T.dj) = s - ls;
    if (A.step == 4) {
        // stage 5's partial sum at this face, which the stage-5 flux launch reads
        // from the stage-5 planes (in its left-to-right order)
        double q = A.b5[0] * get(KH[0], kp, T.di, T.dj);
        for (int k = 1; k < 5; k++) q = q + A.b5[k] * get(KH[k], kp, T.di, T.dj);
        at(KH[5], kp, T.di, T.dj) = q;
    }
}

template <int R>
__device__ __forceinline__ void program_task(const ProgArgs &A, const double *E, const int t) {
    const vx_cell_task T = A.task[t];
    if (T.op >= VX_OP_C_SAME && T.op <= VX_OP_C_UPDATE) {
        c_task<R>(A, T);
        return;
    }
    if (T.op == VX_OP_FLUX_X || T.op == VX_OP_FLUX_P) {
        flux_task<R>(A, E, T);
        return;
    }
    if (T.op == VX_OP_CORRECT) {
        const vx_patch &D = A.pt[T.dst];
        at(D.f[1], D.pitch, T.di, T.dj) = corrected_cell(D, T.di, T.dj);
        return;
    }
    const vx_patch &D = A.pt[T.dst];
    const int rs = A.mode == VX_MODE_TRANSFER ? 1 : A.slot;
    double v = 0.0;
    if (T.op == VX_OP_COPY) {
        const vx_patch &S = A.pt[T.src];
        v = get(S.f[rs], S.pitch, T.si, T.sj);
    } else if (T.op == VX_OP_COARSE) {
        const vx_patch &S = A.pt[T.src];
        v = coarse_element<R>(A, S, S.f[rs], T.si, T.sj, T.aux);
    } else if (T.op == VX_OP_RESTRICT) {
        const vx_patch &S = A.pt[T.src];
        v = finer_mean<R>(A, S, S.f[rs], T.si, T.sj);
    }
    if (A.mode == VX_MODE_TRANSFER) {
        at(D.f[0], D.pitch, T.di, T.dj) = v;
        at(D.f[1], D.pitch, T.di, T.dj) = v;
    } else {
        at(D.f[A.slot], D.pitch, T.di, T.dj) = v;
    }
}

// One batch of a program: one thread per task.
template <int R>
__global__ __launch_bounds__(256) void k_program(ProgArgs A, const double *E) {
    const int t = A.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.end) return;
    program_task<R>(A, E, t);
}

// One batch with rectangles [r0, r1): workgroup b < nrblk covers 256 cells of
// the rectangle whose workgroups it falls in (each cell is the RESTRICT task of
// program_task); the workgroups after those run the batch's cell tasks
// [A.begin, A.end), in the same launch.
template <int R>
__global__ __launch_bounds__(256) void k_restrict_rects(ProgArgs A, const double *E, const vx_restrict_rect *rects,
                                                        int r0, int r1, int nrblk) {
    const int b = (int)blockIdx.x;
    if (b >= nrblk) {
        const int t = A.begin + (b - nrblk) * 256 + (int)threadIdx.x;
        if (t < A.end) program_task<R>(A, E, t);
        return;
    }
    int k = r0;
    while (k + 1 < r1 && rects[k + 1].blk0 <= b) k++;  // uniform: a few rectangles per batch
    const vx_restrict_rect Q = rects[k];
    const int c = (b - Q.blk0) * 256 + (int)threadIdx.x;
    if (c >= Q.ni * Q.nj) return;
    const int di = Q.i0 + c / Q.nj, dj = Q.j0 + c % Q.nj;
    const vx_patch &D = A.pt[Q.dst], &S = A.pt[Q.src];
    const int rs = A.mode == VX_MODE_TRANSFER ? 1 : A.slot;
    const double v = finer_mean<R>(A, S, S.f[rs], di + Q.oi, dj + Q.oj);
    if (A.mode == VX_MODE_TRANSFER) {
        at(D.f[0], D.pitch, di, dj) = v;
        at(D.f[1], D.pitch, di, dj) = v;
    } else {
        at(D.f[A.slot], D.pitch, di, dj) = v;
    }
}

// A run of consecutive small batches of a program in one workgroup: batch b
// (tasks [starts[b], starts[b + 1])) is done by the whole workgroup before
// batch b + 1 starts, the order the batches' separate launches give.  The
// workgroup fence of the barrier makes each batch's writes visible to the
// next batch's reads in this workgroup (one compute unit).
constexpr int SEQ_THREADS = 1024;
template <int R>
__global__ __launch_bounds__(SEQ_THREADS) void k_program_seq(ProgArgs A, const double *E, const int *starts,
                                                             int b0, int b1) {
    for (int b = b0; b < b1; b++) {
        const int lo = starts[b], hi = starts[b + 1];
        for (int t = lo + (int)threadIdx.x; t < hi; t += SEQ_THREADS) program_task<R>(A, E, t);
        __syncthreads();
    }
}

// Rectangle::ErrorEstimate (Rectangle.hpp:128-130) > criteria
__global__ __launch_bounds__(256) void k_flags(vx_patch P, double w0, double w1, double w2, double w3,
                                               double w4, double crit, int lo_x, int hi_x, int lo_p,
                                               int hi_p, uint8_t *flags) {
    const int w = hi_p - lo_p;
    const long n = (long)(hi_x - lo_x) * w;
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i = lo_x + (int)(e / w), j = lo_p + (int)(e % w);
    const double *g = P.f[1];
    const int pt = P.pitch;
    double c = get(g, pt, i, j);
    double err = w0 * fabs(get(g, pt, i + 1, j) - get(g, pt, i - 1, j)) +
                 w1 * fabs(get(g, pt, i, j + 1) - get(g, pt, i, j - 1)) +
                 w2 * fabs(get(g, pt, i + 1, j) - 2 * c + get(g, pt, i - 1, j)) +
                 w3 * fabs(get(g, pt, i, j + 1) - 2 * c + get(g, pt, i, j - 1)) + w4 * fabs(c);
    flags[e] = err > crit ? 1 : 0;
}

// Ordered compaction of the flags: per 1024-cell block, the count (k_flag_count),
// then each flagged cell's index at its block's offset plus its rank in the
// block (wavefront ballots), so the list is in increasing cell index, i.e. the
// reference's row-major scan order (Mesh.cpp:212-295 collects flags that way).
constexpr int FB = 1024;

__global__ __launch_bounds__(FB) void k_flag_count(const uint8_t *flags, long n, int *cnt) {
    const long e = blockIdx.x * (long)FB + threadIdx.x;
    const int c = __syncthreads_count(e < n && flags[e]);
    if (threadIdx.x == 0) cnt[blockIdx.x] = c;
}

__global__ __launch_bounds__(FB) void k_flag_scatter(const uint8_t *flags, long n, const int *off, int w,
                                                     int *list, int *row_sig, int *col_sig) {
    __shared__ int wave_tot[FB / 64];
    const long e = blockIdx.x * (long)FB + threadIdx.x;
    const bool f = e < n && flags[e];
    const unsigned long long m = __ballot(f);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rank = __popcll(m & ((1ull << lane) - 1));
    if (lane == 0) wave_tot[wv] = __popcll(m);
    __syncthreads();
    int base = off[blockIdx.x];
    for (int k = 0; k < wv; k++) base += wave_tot[k];
    if (f) {
        list[base + rank] = (int)e;
        atomicAdd(&row_sig[e / w], 1);  // integer counts: exact in any order
        atomicAdd(&col_sig[e % w], 1);
    }
}

int check_program(vx_ctx *ctx, const vx_patch_table *tb, vx_program *prog, int r);

// One thread per cell of a phase of regrid-transfer rectangles: range k covers
// cells [start[k], start[k+1]) of the launch (binary search on start).
template <int R>
__global__ __launch_bounds__(256) void k_ranges(ProgArgs A, const vx_range_task *rg, const long *start, int n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= start[n]) return;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const vx_range_task G = rg[lo];
    const long local = t - start[lo];
    const vx_patch &D = A.pt[G.dst], &S = A.pt[G.src];
    double v;
    int di, dj;
    if (G.op == VX_OP_COPY) {
        const int a = (int)(local / G.nj), b = (int)(local % G.nj);
        di = G.di0 + a;
        dj = G.dj0 + b;
        v = get(S.f[1], S.pitch, G.si0 + a, G.sj0 + b);
    } else {  // VX_OP_COARSE
        const int rr = R * R;
        const int cell = (int)(local / rr), e = (int)(local % rr);
        const int k = e / R, l = e % R;
        const int ci = G.si0 + cell / G.nj, cj = G.sj0 + cell % G.nj;
        di = (ci - G.di0) * R + k;
        dj = (cj - G.dj0) * R + l;
        v = coarse_element<R>(A, S, S.f[1], R * ci + 1, R * cj + 1, l + R * k);
    }
    at(D.f[0], D.pitch, di, dj) = v;
    at(D.f[1], D.pitch, di, dj) = v;
}

int fill_common(vx_ctx *ctx, const vx_patch_table *table, const vx_program *prog, int r, ProgArgs &A) {
    if (!table || !prog || !table->d) return fail(ctx, VX_EINVAL, "program: null table or program");
    if (r < 2 || r > MAXR) return fail(ctx, VX_ENOTSUP, "program: refinement ratio %d not in [2, %d]", r, MAXR);
    int rc = check_program(ctx, table, const_cast<vx_program *>(prog), r);
    if (rc) return rc;
    A.pt = table->d;
    A.task = prog->d;
    A.r = r;
    A.inv_r2 = 1.0 / std::pow(r, 2);
    A.r2 = std::pow((double)r, 2.0);
    A.inv_rr = (1.0 / (r * r));
    for (int i = 0; i < 3 * MAXR; i++) A.ref[i] = 0.0;
    for (int i = 0; i < r; i++) {  // Rectangle.cpp:91-101
        double tl = -0.5 + i / (double)r;
        double tr = -0.5 + (i + 1.0) / (double)r;
        A.ref[3 * i] = (tl + tr) * 0.5;
        A.ref[3 * i + 1] = (tl * tl + tl * tr + tr * tr) / 3.0 - (1.0 / 12);
        A.ref[3 * i + 2] = (tl * tl * tl + tl * tl * tr + tl * tr * tr + tr * tr * tr) * 0.25;
    }
    return VX_OK;
}

// Host bounds check of every cell a program's tasks touch (ghost ring
// included), done once per (program, table) pair before the first launch.
bool inr(const vx_patch &P, int i, int j) { return i >= -2 && i < P.n_x + 2 && j >= -2 && j < P.n_p + 2; }

int check_program(vx_ctx *ctx, const vx_patch_table *tb, vx_program *prog, int r) {
    if (prog->checked_serial == tb->serial && prog->checked_r == r) return VX_OK;
    const int n = tb->n;
    for (size_t t = 0; t < prog->host.size(); t++) {
        const vx_cell_task &T = prog->host[t];
        bool ok = T.dst >= 0 && T.dst < n;
        const bool uses_src = T.op != VX_OP_ZERO && T.op != VX_OP_CORRECT;
        if (ok && uses_src) ok = T.src >= 0 && T.src < n;
        if (ok) {
            const vx_patch &D = tb->host[T.dst];
            const vx_patch &S = tb->host[uses_src ? T.src : T.dst];
            const int di = T.di, dj = T.dj, si = T.si, sj = T.sj;
            switch (T.op) {
                case VX_OP_ZERO: ok = inr(D, di, dj); break;
                case VX_OP_COPY: ok = inr(D, di, dj) && inr(S, si, sj); break;
                case VX_OP_COARSE: {
                    int ic = si / r - S.x_pos, jc = sj / r - S.p_pos;
                    ok = inr(D, di, dj) && inr(S, ic - 2, jc - 2) && inr(S, ic + 2, jc + 2) && T.aux >= 0 &&
                         T.aux < r * r;
                    break;
                }
                case VX_OP_RESTRICT: {
                    int i0 = si * r - S.x_pos, j0 = sj * r - S.p_pos;
                    ok = inr(D, di, dj) && inr(S, i0, j0) && inr(S, i0 + r - 1, j0 + r - 1);
                    break;
                }
                case VX_OP_C_SAME: case VX_OP_C_UPDATE:
                    ok = T.aux >= 0 && T.aux < 4 && inr(D, di, dj) && inr(S, si, sj) && inr(D, di - 1, dj - 1) &&
                         inr(S, si - 1, sj - 1);
                    break;
                case VX_OP_C_DIFF:
                    ok = T.aux >= 0 && T.aux < 4 && inr(D, di, dj) && inr(D, di - 1, dj - 1) && inr(S, si - 1, sj - 1) &&
                         inr(S, si + r - 1, sj + r - 1);
                    break;
                case VX_OP_FLUX_X: case VX_OP_FLUX_P:
                    ok = inr(D, di, dj) && inr(S, si - 1, sj - 1) && inr(S, si + r - 1, sj + r - 1);
                    break;
                case VX_OP_CORRECT: ok = di >= 0 && di < D.n_x && dj >= 0 && dj < D.n_p; break;
                default: ok = false;
            }
        }
        if (!ok)
            return fail(ctx, VX_EINVAL, "program task %zu (op %d, dst %d, src %d) out of range", t, T.op, T.dst,
                        T.src);
    }
    prog->checked_serial = tb->serial;
    prog->checked_r = r;
    return VX_OK;
}

// Batches of at most SEQ_MAX tasks that follow each other run together in one
// workgroup (k_program_seq): a few passes of one compute unit cost less than
// a launch each.  Larger batches keep a launch of their own over the device.
constexpr int SEQ_MAX = 4 * SEQ_THREADS;

template <int R>
int launch_batches_r(vx_ctx *ctx, const vx_program *prog, ProgArgs &A, const double *E) {
    const int nb = prog->cbatch.empty() ? 0 : (int)prog->cbatch.size() - 1;
    auto has_rects = [&](int b) { return prog->rbatch[b + 1] > prog->rbatch[b]; };
    for (int b = 0; b < nb;) {
        int e = b;
        while (e < nb && !has_rects(e) && prog->cbatch[e + 1] - prog->cbatch[e] <= SEQ_MAX) e++;
        if (e - b >= 2) {  // a run of small batches
            k_program_seq<R><<<1, SEQ_THREADS, 0, ctx->stream>>>(A, E, prog->d_batch, b, e);
            int rc = check_launch(ctx, "k_program_seq");
            if (rc) return rc;
            b = e;
            continue;
        }
        A.begin = prog->cbatch[b];
        A.end = prog->cbatch[b + 1];
        const int n = A.end - A.begin;
        if (has_rects(b)) {  // the rectangles and the batch's cell tasks in one launch
            const int nr = prog->rblocks[b];
            k_restrict_rects<R><<<(unsigned)(nr + (n + 255) / 256), 256, 0, ctx->stream>>>(
                A, E, prog->d_rect, prog->rbatch[b], prog->rbatch[b + 1], nr);
            int rc = check_launch(ctx, "k_restrict_rects");
            if (rc) return rc;
            b++;
            continue;
        }
        b++;
        if (n <= 0) continue;
        k_program<R><<<(n + 255) / 256, 256, 0, ctx->stream>>>(A, E);
        int rc = check_launch(ctx, "k_program");
        if (rc) return rc;
    }
    return VX_OK;
}

int launch_batches(vx_ctx *ctx, const vx_program *prog, ProgArgs &A, const double *E) {
    if (A.r == 2) return launch_batches_r<2>(ctx, prog, A, E);
    if (A.r == 3) return launch_batches_r<3>(ctx, prog, A, E);
    static_assert(MAXR == 4, "one launch_batches_r per admitted ratio");
    return launch_batches_r<4>(ctx, prog, A, E);  // fill_common admits 2..MAXR
}

// Split the RESTRICT tasks of each batch of more than SEQ_MAX tasks into
// rectangles (row runs of consecutive j merged over consecutive i, with the
// same coarse-coordinate offsets) of at least 256 cells; every other task stays
// a cell task.  The tasks of a batch touch disjoint data, so their order within
// the batch is free.
void compact_program(const std::vector<vx_cell_task> &host, const std::vector<int32_t> &batch,
                     std::vector<vx_cell_task> &cells, std::vector<int32_t> &cbatch,
                     std::vector<vx_restrict_rect> &rects, std::vector<int32_t> &rbatch,
                     std::vector<int32_t> &rblocks) {
    const int nb = (int)batch.size() - 1;
    cells.clear();
    rects.clear();
    cbatch.assign(1, 0);
    rbatch.assign(1, 0);
    rblocks.clear();
    for (int b = 0; b < nb; b++) {
        const int lo = batch[b], hi = batch[b + 1];
        std::vector<int> rt;  // RESTRICT tasks of a large batch
        if (hi - lo > SEQ_MAX)
            for (int t = lo; t < hi; t++)
                if (host[t].op == VX_OP_RESTRICT) rt.push_back(t);
        std::sort(rt.begin(), rt.end(), [&](int x, int y) {
            const vx_cell_task &X = host[x], &Y = host[y];
            if (X.dst != Y.dst) return X.dst < Y.dst;
            if (X.src != Y.src) return X.src < Y.src;
            if (X.di != Y.di) return X.di < Y.di;
            return X.dj < Y.dj;
        });
        // row runs, then rectangles of runs with the same columns on consecutive rows
        struct Run {
            int dst, src, i, j0, nj, oi, oj, first;  // first: index into rt
        };
        std::vector<Run> runs;
        for (size_t k = 0; k < rt.size(); k++) {
            const vx_cell_task &T = host[rt[k]];
            const int oi = T.si - T.di, oj = T.sj - T.dj;
            if (!runs.empty()) {
                Run &R = runs.back();
                if (R.dst == T.dst && R.src == T.src && R.i == T.di && R.j0 + R.nj == T.dj && R.oi == oi &&
                    R.oj == oj) {
                    R.nj++;
                    continue;
                }
            }
            runs.push_back({T.dst, T.src, T.di, T.dj, 1, oi, oj, (int)k});
        }
        std::vector<char> in_rect(rt.size(), 0);
        std::vector<vx_restrict_rect> br;
        for (size_t a = 0; a < runs.size();) {
            size_t e = a + 1;
            while (e < runs.size() && runs[e].dst == runs[a].dst && runs[e].src == runs[a].src &&
                   runs[e].i == runs[e - 1].i + 1 && runs[e].j0 == runs[a].j0 && runs[e].nj == runs[a].nj &&
                   runs[e].oi == runs[a].oi && runs[e].oj == runs[a].oj)
                e++;
            const long ncell = (long)(e - a) * runs[a].nj;
            if (ncell >= 256) {
                vx_restrict_rect R{};
                R.dst = runs[a].dst;
                R.src = runs[a].src;
                R.i0 = runs[a].i;
                R.j0 = runs[a].j0;
                R.ni = (int)(e - a);
                R.nj = runs[a].nj;
                R.oi = runs[a].oi;
                R.oj = runs[a].oj;
                br.push_back(R);
                for (size_t q = a; q < e; q++)
                    for (int m = 0; m < runs[q].nj; m++) in_rect[runs[q].first + m] = 1;
            }
            a = e;
        }
        std::vector<char> take(hi - lo, 1);
        for (size_t k = 0; k < rt.size(); k++)
            if (in_rect[k]) take[rt[k] - lo] = 0;
        for (int t = lo; t < hi; t++)
            if (take[t - lo]) cells.push_back(host[t]);
        cbatch.push_back((int32_t)cells.size());
        int blk = 0;
        for (vx_restrict_rect &R : br) {
            R.blk0 = blk;
            blk += (R.ni * R.nj + 255) / 256;
            rects.push_back(R);
        }
        rbatch.push_back((int32_t)rects.size());
        rblocks.push_back(blk);
    }
}

}  // namespace

extern "C" {

int32_t vx_program_batch_sizes(const vx_program *prog, int32_t *sizes, int32_t n) {
    if (!prog || prog->batch.empty()) return 0;
    const int32_t nb = (int32_t)prog->batch.size() - 1;
    if (sizes && n >= nb)
        for (int32_t b = 0; b < nb; b++) sizes[b] = prog->batch[b + 1] - prog->batch[b];
    return nb;
}

int vx_patch_table_create(vx_ctx *ctx, const vx_patch *patches, int32_t n, vx_patch_table **out) {
    if (!out || n < 0 || (n > 0 && !patches)) return fail(ctx, VX_EINVAL, "patch table: bad arguments");
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "patch table while a step graph is recorded");
    vx_patch_table *t = new vx_patch_table;
    t->ctx = ctx;
    t->n = n;
    t->serial = ++g_table_serial;
    t->host.assign(patches, patches + n);
    if (n > 0) {
        hipError_t e = hipMalloc(&t->d, sizeof(vx_patch) * n);
        if (e != hipSuccess) {
            delete t;
            return fail(ctx, VX_ENOMEM, "patch table: %s", hipGetErrorString(e));
        }
        int rc = vx_patch_table_update(ctx, t, patches, n);
        if (rc) {
            vx_patch_table_destroy(t);
            return rc;
        }
    }
    *out = t;
    return VX_OK;
}

int vx_patch_table_update(vx_ctx *ctx, vx_patch_table *t, const vx_patch *patches, int32_t n) {
    if (!t || n != t->n) return fail(ctx, VX_EINVAL, "patch table: size mismatch");
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "patch table upload while a step graph is recorded");
    if (n == 0) return VX_OK;
    for (int k = 0; k < n; k++) {
        const vx_patch &a = t->host[k], &b = patches[k];
        if (a.n_x != b.n_x || a.n_p != b.n_p || a.pitch != b.pitch || a.x_pos != b.x_pos || a.p_pos != b.p_pos)
            return fail(ctx, VX_EINVAL, "patch table: patch %d changed shape", k);
        int rc = validate_patch(ctx, &b);
        if (rc) return rc;
    }
    t->host.assign(patches, patches + n);
    hipError_t e = hipMemcpyAsync(t->d, patches, sizeof(vx_patch) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, VX_EHIP, "patch table upload: %s", hipGetErrorString(e));
    return VX_OK;
}

void vx_patch_table_destroy(vx_patch_table *t) {
    if (!t) return;
    if (t->d) (void)hipFree(t->d);
    if (t->batch_d) (void)hipFree(t->batch_d);
    delete t;
}

int vx_program_create(vx_ctx *ctx, const vx_cell_task *tasks, int32_t n_tasks, const int32_t *batch_start,
                      int32_t n_batches, vx_program **out) {
    if (!out || n_tasks < 0 || n_batches < 0 || (n_tasks > 0 && (!tasks || !batch_start)))
        return fail(ctx, VX_EINVAL, "program: bad arguments");
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "program upload while a step graph is recorded");
    if (n_tasks > 0 && (batch_start[0] != 0 || batch_start[n_batches] != n_tasks))
        return fail(ctx, VX_EINVAL, "program: batches must cover [0, %d)", n_tasks);
    for (int b = 0; b < n_batches; b++)
        if (n_tasks > 0 && batch_start[b] > batch_start[b + 1])
            return fail(ctx, VX_EINVAL, "program: batch starts not ordered");
    vx_program *p = new vx_program;
    p->ctx = ctx;
    if (n_tasks > 0) {
        p->batch.assign(batch_start, batch_start + n_batches + 1);
        p->host.assign(tasks, tasks + n_tasks);
        std::vector<vx_cell_task> cells;
        std::vector<vx_restrict_rect> rects;
        compact_program(p->host, p->batch, cells, p->cbatch, rects, p->rbatch, p->rblocks);
        hipError_t e = hipSuccess;
        if (!cells.empty()) {
            e = hipMalloc(&p->d, sizeof(vx_cell_task) * cells.size());
            if (e == hipSuccess)
                e = hipMemcpy(p->d, cells.data(), sizeof(vx_cell_task) * cells.size(), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess) e = hipMalloc(&p->d_batch, sizeof(int32_t) * (n_batches + 1));
        if (e == hipSuccess)
            e = hipMemcpy(p->d_batch, p->cbatch.data(), sizeof(int32_t) * (n_batches + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess && !rects.empty()) {
            e = hipMalloc(&p->d_rect, sizeof(vx_restrict_rect) * rects.size());
            if (e == hipSuccess)
                e = hipMemcpy(p->d_rect, rects.data(), sizeof(vx_restrict_rect) * rects.size(), hipMemcpyHostToDevice);
        }
        if (e != hipSuccess) {
            if (p->d) (void)hipFree(p->d);
            if (p->d_batch) (void)hipFree(p->d_batch);
            if (p->d_rect) (void)hipFree(p->d_rect);
            delete p;
            return fail(ctx, VX_ENOMEM, "program upload: %s", hipGetErrorString(e));
        }
    }
    *out = p;
    return VX_OK;
}

void vx_program_destroy(vx_program *prog) {
    if (!prog) return;
    if (prog->d) (void)hipFree(prog->d);
    if (prog->d_batch) (void)hipFree(prog->d_batch);
    if (prog->d_rect) (void)hipFree(prog->d_rect);
    delete prog;
}

int vx_transfer_ranges(vx_ctx *ctx, const vx_patch_table *table, const vx_range_task *ranges, int32_t n,
                       int32_t r) {
    if (n == 0) return VX_OK;
    if (!table || !table->d || !ranges || n < 0) return fail(ctx, VX_EINVAL, "transfer ranges: bad arguments");
    if (r < 2 || r > MAXR) return fail(ctx, VX_ENOTSUP, "transfer ranges: refinement ratio %d not in [2, %d]", r, MAXR);
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "transfer ranges while a step graph is recorded");
    // host bounds check of every cell read and written, and the cell prefix
    std::vector<long> start(n + 1, 0);
    for (int k = 0; k < n; k++) {
        const vx_range_task &R = ranges[k];
        bool ok = R.dst >= 0 && R.dst < table->n && R.src >= 0 && R.src < table->n && R.ni > 0 && R.nj > 0;
        long cells = 0;
        if (ok) {
            const vx_patch &D = table->host[R.dst], &S = table->host[R.src];
            if (R.op == VX_OP_COPY) {
                ok = inr(D, R.di0, R.dj0) && inr(D, R.di0 + R.ni - 1, R.dj0 + R.nj - 1) && inr(S, R.si0, R.sj0) &&
                     inr(S, R.si0 + R.ni - 1, R.sj0 + R.nj - 1);
                cells = (long)R.ni * R.nj;
            } else if (R.op == VX_OP_COARSE) {
                const int ic = R.si0 - S.x_pos, jc = R.sj0 - S.p_pos;
                const int fi = (R.si0 - R.di0) * r, fj = (R.sj0 - R.dj0) * r;
                ok = R.si0 >= 0 && R.sj0 >= 0 && inr(S, ic - 2, jc - 2) && inr(S, ic + R.ni + 1, jc + R.nj + 1) &&
                     inr(D, fi, fj) && inr(D, fi + R.ni * r - 1, fj + R.nj * r - 1);
                cells = (long)R.ni * R.nj * r * r;
            } else {
                ok = false;
            }
        }
        if (!ok)
            return fail(ctx, VX_EINVAL, "transfer range %d (op %d, dst %d, src %d) out of range", k, ranges[k].op,
                        ranges[k].dst, ranges[k].src);
        start[k + 1] = start[k] + cells;
    }
    ProgArgs A{};
    A.pt = table->d;
    A.r = r;
    A.inv_r2 = 1.0 / std::pow(r, 2);
    A.r2 = std::pow((double)r, 2.0);
    A.inv_rr = (1.0 / (r * r));
    for (int i = 0; i < 3 * MAXR; i++) A.ref[i] = 0.0;
    for (int i = 0; i < r; i++) {  // Rectangle.cpp:91-101
        double tl = -0.5 + i / (double)r;
        double tr = -0.5 + (i + 1.0) / (double)r;
        A.ref[3 * i] = (tl + tr) * 0.5;
        A.ref[3 * i + 1] = (tl * tl + tl * tr + tr * tr) / 3.0 - (1.0 / 12);
        A.ref[3 * i + 2] = (tl * tl * tl + tl * tl * tr + tl * tr * tr + tr * tr * tr) * 0.25;
    }
    const size_t rb = sizeof(vx_range_task) * n, sb = sizeof(long) * (n + 1);
    char *d = nullptr;
    int rc = check_hip(ctx, hipMalloc((void **)&d, rb + sb), "transfer ranges");
    if (rc) return rc;
    std::vector<char> h(rb + sb);
    std::memcpy(h.data(), ranges, rb);
    std::memcpy(h.data() + rb, start.data(), sb);
    rc = check_hip(ctx, hipMemcpyAsync(d, h.data(), rb + sb, hipMemcpyHostToDevice, ctx->stream), "transfer ranges");
    if (!rc && start[n] > 0) {
        const unsigned g = (unsigned)((start[n] + 255) / 256);
        const vx_range_task *rg = (const vx_range_task *)d;
        const long *st = (const long *)(d + rb);
        if (r == 2) k_ranges<2><<<g, 256, 0, ctx->stream>>>(A, rg, st, n);
        else if (r == 3) k_ranges<3><<<g, 256, 0, ctx->stream>>>(A, rg, st, n);
        else k_ranges<4><<<g, 256, 0, ctx->stream>>>(A, rg, st, n);
        rc = check_launch(ctx, "k_ranges");
    }
    const int rs = check_hip(ctx, hipStreamSynchronize(ctx->stream), "transfer ranges");
    (void)hipFree(d);
    return rc ? rc : rs;
}

int vx_program_run(vx_ctx *ctx, const vx_patch_table *table, const vx_program *prog, int mode, int slot, int r) {
    if (prog && prog->batch.empty()) return VX_OK;
    ProgArgs A{};
    int rc = fill_common(ctx, table, prog, r, A);
    if (rc) return rc;
    if (mode != VX_MODE_GHOST && mode != VX_MODE_TRANSFER) return fail(ctx, VX_EINVAL, "program: bad mode %d", mode);
    if (slot < 0 || slot > 2) return fail(ctx, VX_EINVAL, "program: bad slot %d", slot);
    A.mode = mode;
    A.slot = slot;
    return launch_batches(ctx, prog, A, nullptr);
}

int vx_program_run_fluxes(vx_ctx *ctx, const vx_patch_table *table, const vx_program *prog, int step, double dt,
                          int r, const vx_fields *fl) {
    if (prog && prog->batch.empty()) return VX_OK;
    ProgArgs A{};
    int rc = fill_common(ctx, table, prog, r, A);
    if (rc) return rc;
    if (step < 0 || step > 5 || !fl) return fail(ctx, VX_EINVAL, "flux program: bad step %d", step);
    A.step = step;
    A.ns = step + 1;
    for (int k = 0; k < 6; k++) A.a[k] = 0.0;
    for (int k = 0; k < A.ns; k++) A.a[k] = vx::rk_a[step][k] * dt;
    A.aSum = A.a[0];
    for (int k = 1; k < A.ns; k++) A.aSum = A.aSum + A.a[k];
    for (int k = 0; k < 6; k++) A.b5[k] = vx::rk_a[5][k] * dt;  // as stage 5 forms its a[k]
    A.a2 = fl->a2;
    A.n_f = fl->n_f;
    return launch_batches(ctx, prog, A, fl->E);
}

int vx_flag_cells(vx_ctx *ctx, const vx_patch *p, const double *w, double criteria, int32_t lo_x, int32_t hi_x,
                  int32_t lo_p, int32_t hi_p, uint8_t *flags) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (!w || !flags) return fail(ctx, VX_EINVAL, "flags: null weights or output");
    if (lo_x < 0 || lo_p < 0 || hi_x > p->n_x || hi_p > p->n_p)
        return fail(ctx, VX_EINVAL, "flags: range [%d,%d)x[%d,%d) outside the patch", lo_x, hi_x, lo_p, hi_p);
    if (hi_x <= lo_x || hi_p <= lo_p) return VX_OK;
    long n = (long)(hi_x - lo_x) * (hi_p - lo_p);
    k_flags<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(*p, w[0], w[1], w[2], w[3], w[4], criteria,
                                                                 lo_x, hi_x, lo_p, hi_p, flags);
    return check_launch(ctx, "k_flags");
}


int vx_flag_compact(vx_ctx *ctx, const vx_patch *p, const double *w, double criteria, int32_t lo_x, int32_t hi_x,
                    int32_t lo_p, int32_t hi_p, int32_t *cells, int32_t max_cells, int32_t *count,
                    int32_t *row_sig, int32_t *col_sig) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (!w || !count) return fail(ctx, VX_EINVAL, "flag compaction: null weights or count");
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "flag compaction while a step graph is recorded");
    if (lo_x < 0 || lo_p < 0 || hi_x > p->n_x || hi_p > p->n_p)
        return fail(ctx, VX_EINVAL, "flags: range [%d,%d)x[%d,%d) outside the patch", lo_x, hi_x, lo_p, hi_p);
    *count = 0;
    if (hi_x <= lo_x || hi_p <= lo_p) return VX_OK;
    const int wd = hi_p - lo_p, hx = hi_x - lo_x;
    const long n = (long)hx * wd;
    const int nb = (int)((n + FB - 1) / FB);
    // device scratch: flags | block counts, offsets | list | row, column signatures
    const size_t bytes = (size_t)n + sizeof(int) * (2 * (size_t)nb + (size_t)n + hx + wd) + 64;
    char *d = nullptr;
    if ((rc = check_hip(ctx, hipMalloc((void **)&d, bytes), "flag compaction scratch"))) return rc;
    uint8_t *flags = (uint8_t *)d;
    int *cnt = (int *)(d + ((n + 15) / 16) * 16), *off = cnt + nb, *list = off + nb, *rs = list + n, *cs = rs + hx;
    std::vector<int> hc(nb), ho(nb);
    k_flags<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(*p, w[0], w[1], w[2], w[3], w[4], criteria,
                                                                 lo_x, hi_x, lo_p, hi_p, flags);
    rc = check_launch(ctx, "k_flags");
    if (!rc) {
        k_flag_count<<<(unsigned)nb, FB, 0, ctx->stream>>>(flags, n, cnt);
        rc = check_launch(ctx, "k_flag_count");
    }
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(hc.data(), cnt, sizeof(int) * nb, hipMemcpyDeviceToHost, ctx->stream),
                            "flag counts");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "flag counts");
    long total = 0;
    for (int b = 0; b < nb; b++) {
        ho[b] = (int)total;
        total += hc[b];
    }
    if (!rc && cells && total > max_cells) rc = fail(ctx, VX_EINVAL, "flag compaction: %ld flagged cells > %d", total, max_cells);
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(off, ho.data(), sizeof(int) * nb, hipMemcpyHostToDevice, ctx->stream),
                            "flag offsets");
    if (!rc) rc = check_hip(ctx, hipMemsetAsync(rs, 0, sizeof(int) * (hx + wd), ctx->stream), "signatures");
    if (!rc) {
        k_flag_scatter<<<(unsigned)nb, FB, 0, ctx->stream>>>(flags, n, off, wd, list, rs, cs);
        rc = check_launch(ctx, "k_flag_scatter");
    }
    if (!rc && cells && total > 0)
        rc = check_hip(ctx, hipMemcpyAsync(cells, list, sizeof(int) * total, hipMemcpyDeviceToHost, ctx->stream), "list");
    if (!rc && row_sig)
        rc = check_hip(ctx, hipMemcpyAsync(row_sig, rs, sizeof(int) * hx, hipMemcpyDeviceToHost, ctx->stream), "rows");
    if (!rc && col_sig)
        rc = check_hip(ctx, hipMemcpyAsync(col_sig, cs, sizeof(int) * wd, hipMemcpyDeviceToHost, ctx->stream), "cols");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "flag compaction");
    (void)hipFree(d);
    if (!rc) *count = (int32_t)total;
    return rc;
}

}  // extern "C"
