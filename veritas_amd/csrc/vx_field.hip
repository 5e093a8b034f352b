// 1-D field solver on the finest grid (EMFieldSolver, EMSolver.cpp).
//
// Poisson (EMSolver.cpp:28-88,156-192).  The reference factors the dense N x N
// matrix with LAPACK and solves it at every RK stage, O(N^2) per solve.  Here
// the same matrix is solved to near correct rounding in O(N) by one
// 1024-thread workgroup:
//   rows 1..N-1 of the matrix are the periodic 4th-order stencil
//   (w2, w1, w3, w1, w2) with column 0 removed, i.e. the periodic operator with
//   phi_0 pinned to 0; row 0 then gives phi_0.  The periodic operator factors
//   exactly as L4 = L2 * M with L2 = [-1, 2, -1] and M = [-1, 14, -1] / 12, so
//   an approximate solve is two prefix sums (L2^-1 on zero-mean data) and a
//   33-tap convolution (M^-1, whose entries decay as (7 - 4 sqrt 3)^|k|).
//   Three rounds of iterative refinement with a double-double residual of the
//   reference's own coefficients take the result to the correctly rounded
//   solution (the CPU oracle reaches the same point with a band LU).
//
// Transverse fields (EMSolver.cpp:194-338, 479-619): one thread per cell.
#include <cmath>
#include <vector>

#include "vx_dev.h"
#include "vx_internal.h"

using namespace vxi;
using namespace vxd;

namespace {

constexpr int PT = 1024;  // threads of the Poisson workgroup
constexpr int NG = 17;    // convolution half-width + 1 for M^-1

struct PoissonArgs {
    vx_field_state fs;
    double w1, w2, w3;
    double g[NG];  // M^-1 coefficients, g[|k|]
    double invN;
};

// Work layout (doubles): b | x | r | c | y | tmp | partials
__host__ __device__ inline size_t wk_off(int n, int which) { return (size_t)which * n; }

constexpr int PW = PT / 64;  // wavefronts in the Poisson workgroup

__device__ __forceinline__ double wave_incl_scan(double v) {
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        double t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// exclusive prefix sum of v[0..n) into out[0..n) (out may alias v), in
// tiles of PT * 4 entries: each thread takes 4 consecutive entries of a tile
// (a wavefront reads 2 KB contiguously), the tile's thread totals are scanned
// across the workgroup and a running carry links the tiles.  Fixed order.
__device__ void block_exclusive_scan(const double *v, double *out, int n, double *lds) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    constexpr int PER = 4, TILE = PT * PER;
    double carry = 0.0;
    for (int base = 0; base < n; base += TILE) {
        const int a = base + t * PER;
        double e[PER];
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            e[k] = (a + k < n) ? v[a + k] : 0.0;
            s += e[k];
        }
        double inc = wave_incl_scan(s);
        if (lane == 63) lds[w] = inc;
        __syncthreads();
        if (t == 0) {
            double run = 0.0;
            for (int k = 0; k < PW; k++) {
                double tot = lds[k];
                lds[PW + k] = run;
                run += tot;
            }
            lds[2 * PW] = run;
        }
        __syncthreads();
        double run = carry + (lds[PW + w] + (inc - s));
#pragma unroll
        for (int k = 0; k < PER; k++)
            if (a + k < n) {
                out[a + k] = run;
                run += e[k];
            }
        carry += lds[2 * PW];
        __syncthreads();
    }
}

// block sum of v[lo..hi), deterministic order (coalesced strided partials)
__device__ double block_sum(const double *v, int lo, int hi, double *lds) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    double s = 0.0;
    for (int i = lo + t; i < hi; i += PT) s += v[i];
    s = wave_sum(s);
    if (lane == 0) lds[w] = s;
    __syncthreads();
    if (t == 0) {
        double tot = 0.0;
        for (int k = 0; k < PW; k++) tot += lds[k];
        lds[2 * PW] = tot;
    }
    __syncthreads();
    double r = lds[2 * PW];
    __syncthreads();
    return r;
}

// Approximate solve of A c = r (A as EMSolver.cpp:41-67 assembles it).
__device__ void approx_solve(const PoissonArgs &P, const double *r, double *c, double *y,
                             double *tmp, double *lds) {
    const int N = P.fs.n_f;
    const int t = threadIdx.x;
    // zero-mean right-hand side of the periodic system: r'_0 = -sum_{i>=1} r_i
    double tail = block_sum(r, 1, N, lds);
    for (int i = t; i < N; i += PT) tmp[i] = (i == 0) ? -tail : r[i];
    __syncthreads();
    // L2 y = r' with L2 = [-1, 2, -1]: the differences d_i = y_i - y_{i-1}
    // obey d_i = d_0 - S_i with S_i = sum_{k<i} r'_k, and periodicity fixes
    // d_0 = mean(S).  With y_0 = 0, y_i = sum_{k=1..i} d_k.
    block_exclusive_scan(tmp, tmp, N, lds);  // tmp = S
    double d0 = block_sum(tmp, 0, N, lds) * P.invN;
    for (int i = t; i < N; i += PT) y[i] = (i == 0) ? 0.0 : d0 - tmp[i];  // D_i = d_i (D_0 = 0)
    __syncthreads();
    block_exclusive_scan(y, tmp, N, lds);  // tmp_i = sum_{k<i} D_k
    for (int i = t; i < N; i += PT) tmp[i] = tmp[i] + y[i];  // y_i = sum_{k<=i} D_k
    __syncthreads();
    // c' = M^-1 y (periodic 33-tap convolution)
    for (int i = t; i < N; i += PT) {
        double s = P.g[0] * tmp[i];
        for (int k = 1; k < NG; k++) {
            int ip = i + k, im = i - k;
            ip -= ip >= N ? N : 0;
            im += im < 0 ? N : 0;
            s += P.g[k] * (tmp[ip] + tmp[im]);
        }
        c[i] = s;
    }
    __syncthreads();
    double c0 = c[0];
    __syncthreads();
    for (int i = t; i < N; i += PT)
        if (i > 0) c[i] = c[i] - c0;
    __syncthreads();
    if (t == 0)
        c[0] = r[0] - (((P.w1 * c[1] + P.w2 * c[2]) + P.w2 * c[N - 2]) + P.w1 * c[N - 1]);
    __syncthreads();
}

// res = b - A x in double-double, rounded once (term order as the oracle's)
__device__ void residual(const PoissonArgs &P, const double *b, const double *x, double *res) {
    const int N = P.fs.n_f;
    for (int i = threadIdx.x; i < N; i += PT) {
        double sh = b[i], sl = 0.0;
        auto sub = [&](double a, double xv) {
            double p = a * xv;
            double e = fma(a, xv, -p);
            double tt = sh - p;
            double bv = tt - sh;
            double err = (sh - (tt - bv)) + (-p - bv);
            sl = sl + (err - e);
            sh = tt;
        };
        if (i == 0) {
            sub(1.0, x[0]); sub(P.w1, x[1]); sub(P.w2, x[2]); sub(P.w2, x[N - 2]); sub(P.w1, x[N - 1]);
        } else {
            const double cf[5] = {P.w2, P.w1, P.w3, P.w1, P.w2};
            for (int d = -2; d <= 2; d++) {
                int j = i + d;
                j = j < 0 ? j + N : (j >= N ? j - N : j);
                if (j == 0) continue;
                sub(cf[d + 2], x[j]);
            }
        }
        res[i] = sh + sl;
    }
    __syncthreads();
}

// EMFieldSolver::GetEfield (EMSolver.cpp:137-154)
__device__ __forceinline__ double efield(const double *phi, int N, double fc, double ex0, int i) {
    int ip1 = i + 1, im1 = i - 1, ip2 = i + 2, im2 = i - 2;
    ip1 = (ip1 > -1) ? ip1 : (ip1 + N);
    im1 = (im1 > -1) ? im1 : (im1 + N);
    ip2 = (ip2 > -1) ? ip2 : (ip2 + N);
    im2 = (im2 > -1) ? im2 : (im2 + N);
    ip1 = (ip1 < N) ? ip1 : (ip1 - N);
    im1 = (im1 < N) ? im1 : (im1 - N);
    ip2 = (ip2 < N) ? ip2 : (ip2 - N);
    im2 = (im2 < N) ? im2 : (im2 - N);
    return -fc * (8 * (phi[ip1] - phi[im1]) - phi[ip2] + phi[im2]) + ex0;
}

__global__ __launch_bounds__(PT) void k_poisson(PoissonArgs P) {
    const vx_field_state &F = P.fs;
    const int N = F.n_f;
    const int t = threadIdx.x;
    __shared__ double lds[PT];
    double *b = F.work + wk_off(N, 0);
    double *x = F.work + wk_off(N, 1);
    double *r = F.work + wk_off(N, 2);
    double *c = F.work + wk_off(N, 3);
    double *y = F.work + wk_off(N, 4);
    double *tmp = F.work + wk_off(N, 5);
    // EMSolver.cpp:157-167: PHI = eps0_inv (charge + neutral), then * dx^2
    const double te = vx_eps0_inv;
    for (int i = t; i < N; i += PT) {
        double v = te * (F.charge[i] + F.neutral[i]);
        v *= F.dx2;
        b[i] = v;
    }
    __syncthreads();
    approx_solve(P, b, x, y, tmp, lds);
    for (int it = 0; it < 3; it++) {
        residual(P, b, x, r);
        approx_solve(P, r, c, y, tmp, lds);
        for (int i = t; i < N; i += PT) x[i] = x[i] + c[i];
        __syncthreads();
    }
    for (int i = t; i < N; i += PT) F.phi[i] = x[i];
    __syncthreads();
    // EMSolver.cpp:191: Ex0 += -(E(-1) + E(0)) * 0.5 with the previous Ex0
    double ex0 = *F.Ex0;
    double e_m1 = efield(F.phi, N, F.field_coef, ex0, -1);
    double e_0 = efield(F.phi, N, F.field_coef, ex0, 0);
    double ex0n = ex0 + -(e_m1 + e_0) * 0.5;
    __syncthreads();
    if (t == 0) *F.Ex0 = ex0n;
    for (int i = t; i < N; i += PT) F.E[i] = efield(F.phi, N, F.field_coef, ex0n, i);
}

// Register/LDS variant for n_f <= 4 * PT: thread t owns the CH consecutive
// entries [t*CH, t*CH + CH); neighbours (residual, convolution) come from LDS.
constexpr int SMALL_MAX = 4 * PT;

// exclusive scan of per-thread totals; returns this thread's offset
__device__ double block_offsets(double tot, double *part) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    double inc = wave_incl_scan(tot);
    if (lane == 63) part[w] = inc;
    __syncthreads();
    if (t == 0) {
        double run = 0.0;
        for (int k = 0; k < PW; k++) {
            double v = part[k];
            part[PW + k] = run;
            run += v;
        }
    }
    __syncthreads();
    double off = part[PW + w] + (inc - tot);
    __syncthreads();
    return off;
}

__device__ double block_total(double v, double *part) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    v = wave_sum(v);
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (t == 0) {
        double tot = 0.0;
        for (int k = 0; k < PW; k++) tot += part[k];
        part[2 * PW] = tot;
    }
    __syncthreads();
    double r = part[2 * PW];
    __syncthreads();
    return r;
}

template <int CH>
__device__ void approx_small(const PoissonArgs &P, const double *in, double *out, double *yb,
                             double *part) {
    const int N = P.fs.n_f;
    const int base = threadIdx.x * CH;
    double loc = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k >= 1 && base + k < N) loc += in[k];
    double tail = block_total(loc, part);
    double rp[CH], S[CH];
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        rp[k] = (i == 0) ? -tail : (i < N ? in[k] : 0.0);
        tot += rp[k];
    }
    double run = block_offsets(tot, part);
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        S[k] = run;
        run += rp[k];
        if (base + k < N) ssum += S[k];
    }
    double d0 = block_total(ssum, part) * P.invN;
    double D[CH];
    tot = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        D[k] = (i == 0 || i >= N) ? 0.0 : d0 - S[k];
        tot += D[k];
    }
    run = block_offsets(tot, part);
#pragma unroll
    for (int k = 0; k < CH; k++) {
        run += D[k];
        if (base + k < N) yb[base + k] = run;
    }
    __syncthreads();
    double cp[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        double v = 0.0;
        if (i < N) {
            v = P.g[0] * yb[i];
            for (int m = 1; m < NG; m++) {
                int ip = i + m, im = i - m;
                ip -= ip >= N ? N : 0;
                im += im < 0 ? N : 0;
                v += P.g[m] * (yb[ip] + yb[im]);
            }
        }
        cp[k] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) part[0] = cp[0];
    __syncthreads();
    double c0 = part[0];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        out[k] = (i > 0) ? cp[k] - c0 : cp[k];
        if (i < N) yb[i] = out[k];
    }
    __syncthreads();
    if (threadIdx.x == 0)
        out[0] = in[0] - (((P.w1 * yb[1] + P.w2 * yb[2]) + P.w2 * yb[N - 2]) + P.w1 * yb[N - 1]);
    __syncthreads();
}

template <int CH>
__device__ void residual_small(const PoissonArgs &P, const double *b, const double *x, double *r,
                               double *xb) {
    const int N = P.fs.n_f;
    const int base = threadIdx.x * CH;
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k < N) xb[base + k] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        if (i >= N) {
            r[k] = 0.0;
            continue;
        }
        double sh = b[k], sl = 0.0;
        auto sub = [&](double a, double xv) {
            double p = a * xv;
            double e = fma(a, xv, -p);
            double tt = sh - p;
            double bv = tt - sh;
            double err = (sh - (tt - bv)) + (-p - bv);
            sl = sl + (err - e);
            sh = tt;
        };
        if (i == 0) {
            sub(1.0, xb[0]); sub(P.w1, xb[1]); sub(P.w2, xb[2]); sub(P.w2, xb[N - 2]); sub(P.w1, xb[N - 1]);
        } else {
            const double cf[5] = {P.w2, P.w1, P.w3, P.w1, P.w2};
            for (int d = -2; d <= 2; d++) {
                int jj = i + d;
                jj = jj < 0 ? jj + N : (jj >= N ? jj - N : jj);
                if (jj == 0) continue;
                sub(cf[d + 2], xb[jj]);
            }
        }
        r[k] = sh + sl;
    }
    __syncthreads();
}

template <int CH>
__global__ __launch_bounds__(PT) void k_poisson_small(PoissonArgs P) {
    const vx_field_state &F = P.fs;
    const int N = F.n_f;
    const int t = threadIdx.x, base = t * CH;
    __shared__ double wb[SMALL_MAX];  // neighbour exchange (residual and convolution phases never overlap)
    __shared__ double part[2 * PW + 2];
    double *xb = wb, *yb = wb;
    double b[CH], x[CH], r[CH], c[CH];
    const double te = vx_eps0_inv;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        double v = 0.0;
        if (i < N) {
            v = te * (F.charge[i] + F.neutral[i]);
            v *= F.dx2;
        }
        b[k] = v;
    }
    approx_small<CH>(P, b, x, yb, part);
    for (int it = 0; it < 3; it++) {
        residual_small<CH>(P, b, x, r, xb);
        approx_small<CH>(P, r, c, yb, part);
#pragma unroll
        for (int k = 0; k < CH; k++) x[k] = x[k] + c[k];
    }
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k < N) {
            xb[base + k] = x[k];
            F.phi[base + k] = x[k];
        }
    __syncthreads();
    double ex0 = *F.Ex0;
    double e_m1 = efield(xb, N, F.field_coef, ex0, -1);
    double e_0 = efield(xb, N, F.field_coef, ex0, 0);
    double ex0n = ex0 + -(e_m1 + e_0) * 0.5;
    __syncthreads();
    if (t == 0) *F.Ex0 = ex0n;
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k < N) F.E[base + k] = efield(xb, N, F.field_coef, ex0n, base + k);
}

// ---------------------------------------------------------------- Maxwell
struct MaxArgs {
    vx_field_state fs;
    int step;
    double by0, bz0;
    double dx_inv;
    double a[6], b[6];
};

__device__ __forceinline__ int fidx(int ne, int i, int slot) { return slot * ne + i; }

// EMSolver.cpp:479-553
__global__ void k_maxwell_rhs(MaxArgs A) {
    const vx_field_state &F = A.fs;
    const int ne = F.n_e, npre = F.n_prepad, N = F.n_f;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    const int s = fidx(ne, i, A.step + 2);
    double *By = F.By, *Bz = F.Bz, *Ey = F.Ey, *Ez = F.Ez, *Ay = F.Ay, *Az = F.Az;
    const double dx_inv = A.dx_inv;
    double rBy = 0, rBz = 0, rEy = 0, rEz = 0, rAy = 0, rAz = 0;
    int i1 = fidx(ne, i, 1);
    if (i == 0) {
        int ip1 = fidx(ne, i + 1, 1);
        rBy = dx_inv * (Ez[ip1] - Ez[i1]);
        rBz = -dx_inv * (Ey[ip1] - Ey[i1]);
        rEy = -dx_inv * (Bz[i1] - A.bz0) * vx_eps0_inv * vx_mu_inv;
        rEz = dx_inv * (By[i1] - A.by0) * vx_eps0_inv * vx_mu_inv;
        rAy = -Ey[i1];
        rAz = -Ez[i1];
    } else if (i < npre || (i >= npre + N && i < ne - 2)) {
        int ip1 = fidx(ne, i + 1, 1), in1 = fidx(ne, i - 1, 1);
        rBy = dx_inv * (Ez[ip1] - Ez[i1]);
        rBz = -dx_inv * (Ey[ip1] - Ey[i1]);
        rEy = -dx_inv * (Bz[i1] - Bz[in1]) * vx_eps0_inv * vx_mu_inv;
        rEz = dx_inv * (By[i1] - By[in1]) * vx_eps0_inv * vx_mu_inv;
        rAy = -Ey[i1];
        rAz = -Ez[i1];
    } else if (i >= npre && i < npre + N) {
        const double c1 = -1.0 / 24;
        const double c2 = 9.0 / 8.0;
        int ip1 = fidx(ne, i + 1, 1), in1 = fidx(ne, i - 1, 1);
        int ip2 = fidx(ne, i + 2, 1), in2 = fidx(ne, i - 2, 1);
        double Jc = F.J[i - npre];
        rBy = dx_inv * (c1 * (Ez[ip2] - Ez[in1]) + c2 * (Ez[ip1] - Ez[i1]));
        rBz = -dx_inv * (c1 * (Ey[ip2] - Ey[in1]) + c2 * (Ey[ip1] - Ey[i1]));
        rEy = -dx_inv * (c1 * (Bz[ip1] - Bz[in2]) + c2 * (Bz[i1] - Bz[in1])) * vx_eps0_inv * vx_mu_inv -
              vx_eps0_inv * Jc * Ay[i1];
        rEz = dx_inv * (c1 * (By[ip1] - By[in2]) + c2 * (By[i1] - By[in1])) * vx_eps0_inv * vx_mu_inv -
              vx_eps0_inv * Jc * Az[i1];
        rAy = -Ey[i1];
        rAz = -Ez[i1];
    }
    By[s] = rBy; Bz[s] = rBz; Ey[s] = rEy; Ez[s] = rEz; Ay[s] = rAy; Az[s] = rAz;
}

// EMSolver.cpp:204-338
__global__ void k_maxwell_update(MaxArgs A) {
    const vx_field_state &F = A.fs;
    const int ne = F.n_e;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    double *Fs[6] = {F.By, F.Bz, F.Ey, F.Ez, F.Ay, F.Az};
    if (A.step <= 4) {
        for (int v = 0; v < 6; v++) {
            double *y = Fs[v];
            double t = y[fidx(ne, i, 0)] + A.a[0] * y[fidx(ne, i, 2)];
            for (int k = 1; k <= A.step; k++) t = t + A.a[k] * y[fidx(ne, i, k + 2)];
            y[fidx(ne, i, 1)] = t;
        }
    } else {
        for (int v = 0; v < 6; v++) {
            double *y = Fs[v];
            double t = y[fidx(ne, i, 1)] + A.b[0] * y[fidx(ne, i, 2)];
            for (int k = 1; k < 6; k++) t = t + A.b[k] * y[fidx(ne, i, k + 2)];
            y[fidx(ne, i, 0)] = t;
            y[fidx(ne, i, 1)] = t;
        }
    }
}

// EMSolver.cpp:555-619; a_squared[n_f] is never written by the reference
__global__ void k_faces(vx_field_state F) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.n_f) return;
    const int ne = F.n_e, np_ = F.n_prepad;
    const double *Ay = F.Ay, *Az = F.Az;
    double ay = face_weno(Ay[fidx(ne, np_ + i - 2, 1)], Ay[fidx(ne, np_ + i - 1, 1)],
                          Ay[fidx(ne, np_ + i, 1)], Ay[fidx(ne, np_ + i + 1, 1)]);
    double az = face_weno(Az[fidx(ne, np_ + i - 2, 1)], Az[fidx(ne, np_ + i - 1, 1)],
                          Az[fidx(ne, np_ + i, 1)], Az[fidx(ne, np_ + i + 1, 1)]);
    F.a2[i] = (ay * ay) + (az * az);
}

__global__ void k_neutralize(vx_field_state F) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.n_f) return;
    F.J[i] = 0.0;
    F.neutral[i] = -F.charge[i];
}

int validate_fs(vx_ctx *ctx, const vx_field_state *fs) {
    if (!fs) return fail(ctx, VX_EINVAL, "null field state");
    if (fs->n_f < 2 * NG || fs->n_prepad < 2 || fs->n_postpad < 2 ||
        fs->n_e != fs->n_f + fs->n_prepad + fs->n_postpad)
        return fail(ctx, VX_EINVAL, "bad field sizes n_f=%d pre=%d post=%d n_e=%d", fs->n_f,
                    fs->n_prepad, fs->n_postpad, fs->n_e);
    return VX_OK;
}

}  // namespace

extern "C" {

size_t vx_field_work_elems(int32_t n_f) { return (size_t)6 * n_f; }

int vx_poisson(vx_ctx *ctx, const vx_field_state *fs) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    PoissonArgs P;
    P.fs = *fs;
    P.w1 = -16.0 / 30.0 / 0.4;  // EMSolver.cpp:28
    P.w2 = 1.0 / 30.0 / 0.4;
    P.w3 = 1 / 0.4;
    double rho = 7.0 - std::sqrt(48.0);
    double C = 12.0 / (14.0 - 2.0 * rho);
    double pw = 1.0;
    for (int k = 0; k < NG; k++) {
        P.g[k] = C * pw;
        pw *= rho;
    }
    P.invN = 1.0 / fs->n_f;
    LaunchTimer lt(ctx, "k_poisson");
    int ch = (fs->n_f + PT - 1) / PT;
    if (fs->n_f <= SMALL_MAX) {
        if (ch <= 1) k_poisson_small<1><<<1, PT, 0, ctx->stream>>>(P);
        else if (ch <= 2) k_poisson_small<2><<<1, PT, 0, ctx->stream>>>(P);
        else k_poisson_small<4><<<1, PT, 0, ctx->stream>>>(P);
    } else {
        k_poisson<<<1, PT, 0, ctx->stream>>>(P);
    }
    return check_launch(ctx, "k_poisson");
}

int vx_maxwell_stage(vx_ctx *ctx, const vx_field_state *fs, int step, double dt, double by0,
                     double bz0) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    if (step < 0 || step > 5) return fail(ctx, VX_EINVAL, "maxwell: bad step %d", step);
    MaxArgs A;
    A.fs = *fs;
    A.step = step;
    A.by0 = by0;
    A.bz0 = bz0;
    A.dx_inv = 1 / fs->dx_f;
    for (int k = 0; k < 6; k++) A.a[k] = A.b[k] = 0.0;
    if (step <= 4) {
        for (int k = 0; k <= step; k++) A.a[k] = vx::rk_a[step][k] * dt;
    } else {
        double a[6];
        for (int k = 0; k < 5; k++) a[k] = vx::rk_a[4][k] * dt;
        a[5] = 0;
        for (int k = 0; k < 6; k++) A.b[k] = vx::rk_a[5][k] * dt - a[k];
    }
    unsigned g = (unsigned)((fs->n_e + 255) / 256);
    k_maxwell_rhs<<<g, 256, 0, ctx->stream>>>(A);
    if ((rc = check_launch(ctx, "k_maxwell_rhs"))) return rc;
    k_maxwell_update<<<g, 256, 0, ctx->stream>>>(A);
    if ((rc = check_launch(ctx, "k_maxwell_update"))) return rc;
    k_faces<<<(unsigned)((fs->n_f + 255) / 256), 256, 0, ctx->stream>>>(*fs);
    return check_launch(ctx, "k_faces");
}

int vx_neutralize(vx_ctx *ctx, const vx_field_state *fs) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    k_neutralize<<<(unsigned)((fs->n_f + 255) / 256), 256, 0, ctx->stream>>>(*fs);
    return check_launch(ctx, "k_neutralize");
}

// EMSolver.cpp:631-664, evaluated on the host from downloaded arrays (the
// reference indexes the slot-0 fields without the n_prepad offset).
int vx_cfl_bound(vx_ctx *ctx, const vx_field_state *fs, int n_species, const double *q,
                 const double *m, const double *dp_finest, double *out) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    int N = fs->n_f;
    std::vector<double> Ay(N), Az(N), By(N), Bz(N), E(N);
    if ((rc = vx_memcpy_d2h(ctx, Ay.data(), fs->Ay, sizeof(double) * N))) return rc;
    if ((rc = vx_memcpy_d2h(ctx, Az.data(), fs->Az, sizeof(double) * N))) return rc;
    if ((rc = vx_memcpy_d2h(ctx, By.data(), fs->By, sizeof(double) * N))) return rc;
    if ((rc = vx_memcpy_d2h(ctx, Bz.data(), fs->Bz, sizeof(double) * N))) return rc;
    if ((rc = vx_memcpy_d2h(ctx, E.data(), fs->E, sizeof(double) * N))) return rc;
    std::vector<double> dps(n_species);
    double dpsMax(0.0);
    for (int i = 0; i < n_species; i++) {
        dps[i] = 1 / dp_finest[i];
        dpsMax = std::max(dpsMax, std::fabs(q[i]) * dps[i]);
    }
    double pc(0.0);
    for (int i = 0; i < N; i++) {
        double Azv = Az[i], Ayv = Ay[i];
        double As = Ayv * Ayv + Azv * Azv;
        double temp(0.0);
        for (int j = 0; j < n_species; j++)
            temp = std::max(temp, std::fabs(q[j]) / m[j] / std::sqrt(1 + As / ((m[j] * vx_cs) * (m[j] * vx_cs))) * dps[j]);
        pc = std::max(pc, temp * std::fabs(Ayv * Bz[i] - Azv * By[i]) + dpsMax * std::fabs(E[i]));
    }
    *out = 1.0 / std::max((vx_cs / fs->dx_f + pc), 1e-40);
    return VX_OK;
}

}  // extern "C"
