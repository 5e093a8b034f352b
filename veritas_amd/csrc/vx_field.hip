// 1-D field solver on the finest grid (EMFieldSolver, EMSolver.cpp).
//
// Poisson (EMSolver.cpp:28-88,156-192).  The reference factors the dense N x N
// matrix with LAPACK and solves it at every RK stage, O(N^2) per solve.  Here
// the same matrix is solved to near correct rounding in O(N) (one 1024-thread
// workgroup with the vector in registers up to N = 4096, a sequence of
// multi-workgroup launches above):
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
#include <algorithm>
#include <cmath>
#include <mutex>
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

// ------------------------------------------------ multi-workgroup solve, n_f > 4096
// Workgroup b owns the CHUNK consecutive entries [b*CHUNK, b*CHUNK + CHUNK).
// One approximate solve of A c = v (the same steps as the single-workgroup
// variant below) is three launches:
//   k_pm_scan_s  S_i = sum_{k<i} v'_k with v'_0 = -sum_{k>=1} v_k, and per-chunk sums of S;
//   k_pm_scan_y  d0 = mean(S), D_i = d0 - S_i (D_0 = 0), y_i = sum_{k<=i} D_k;
//   k_pm_conv    c = M^-1 y - (M^-1 y)_0 with c_0 from row 0, x (+)= c, then either the
//                next residual r = b - A x (double-double) with its per-chunk sums, or
//                phi, Ex0 and E.
// A sum across workgroups is formed from per-chunk partials that every workgroup
// reads and adds in the same order, so all of them hold the same bits.  k_pm_conv
// recomputes x on a 2-entry halo around its chunk (for the residual and E
// stencils) and at the six entries around 0 (row 0 and Ex0) with the owner's
// operations, so every copy of an entry has the owner's bits.  x and r alternate
// between two buffers: a launch reads one and writes the other.
constexpr int MT = 256, MW = MT / 64, MPER = 4, CHUNK = MT * MPER;
constexpr int HALO = 2;
constexpr int CW = NG - 1 + HALO;  // y window margin: convolution reach + halo
constexpr int NSPEC = 6;           // entries N-3, N-2, N-1, 0, 1, 2

__host__ __device__ constexpr int pm_chunks(int n) { return (n + CHUNK - 1) / CHUNK; }

// Work layout (doubles): b | xa | xb | ra | rb | S | y | Pa[nb] | Pb[nb] | Ex0 on entry
struct PmWork {
    double *b, *xa, *xb, *ra, *rb, *S, *y, *Pa, *Pb, *ex0;
};
__host__ __device__ inline PmWork pm_work(double *w, int n) {
    const size_t N = n;
    const int nb = pm_chunks(n);
    PmWork W;
    W.b = w;
    W.xa = w + N;
    W.xb = w + 2 * N;
    W.ra = w + 3 * N;
    W.rb = w + 4 * N;
    W.S = w + 5 * N;
    W.y = w + 6 * N;
    W.Pa = w + 7 * N;
    W.Pb = W.Pa + nb;
    W.ex0 = W.Pb + nb;
    return W;
}

struct PoissonMB {
    vx_field_state fs;
    double w1, w2, w3;
    double g[NG];
    double invN;
    int nb;
    PmWork W;
    const double *in;    // right-hand side of this solve (b, then residuals)
    const double *x_in;  // x before this solve's correction
    double *x_out;       // x after it
    double *r_out;       // next residual
};

__device__ __forceinline__ int wrap(int u, int N) { return u < 0 ? u + N : (u >= N ? u - N : u); }

// sum of v[lo..hi) in a fixed order; every lane of every wave gets the same bits
__device__ __forceinline__ double wave_range_sum(const double *v, int lo, int hi) {
    double s = 0.0;
    for (int k = lo + (threadIdx.x & 63); k < hi; k += 64) s += v[k];
    return wave_sum(s);
}

__device__ __forceinline__ double mb_sum(double v, double *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    __syncthreads();
    return r;
}

// exclusive prefix of the per-thread totals over the workgroup
__device__ __forceinline__ double mb_offset(double tot, double *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double inc = wave_incl_scan(tot);
    double exc = __shfl_up(inc, 1, 64);
    if (lane == 0) exc = 0.0;
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    double off = 0.0;
    for (int k = 0; k < w; k++) off += lds[k];
    __syncthreads();
    return off + exc;
}

// (M^-1 y) at window position c: the periodic 33-tap convolution
__device__ __forceinline__ double conv_at(const double *w, int c, const double *g) {
    double s = g[0] * w[c];
    for (int k = 1; k < NG; k++) s += g[k] * (w[c + k] + w[c - k]);
    return s;
}

// b = dx^2 eps0_inv (charge + neutral) (EMSolver.cpp:157-167), its chunk sums
// over entries >= 1, and the entry value of Ex0
__global__ __launch_bounds__(MT) void k_pm_rhs(PoissonMB P) {
    __shared__ double lds[MW];
    const vx_field_state &F = P.fs;
    const int N = F.n_f, base = blockIdx.x * CHUNK + threadIdx.x * MPER;
    const double te = vx_eps0_inv;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < MPER; k++) {
        const int i = base + k;
        if (i < N) {
            double v = te * (F.charge[i] + F.neutral[i]);
            v *= F.dx2;
            P.W.b[i] = v;
            if (i >= 1) s += v;
        }
    }
    const double tot = mb_sum(s, lds);
    if (threadIdx.x == 0) {
        P.W.Pa[blockIdx.x] = tot;
        if (blockIdx.x == 0) *P.W.ex0 = *F.Ex0;
    }
}

__global__ __launch_bounds__(MT) void k_pm_scan_s(PoissonMB P) {
    __shared__ double lds[MW];
    const int N = P.fs.n_f, blk = blockIdx.x, base = blk * CHUNK + threadIdx.x * MPER;
    const double tail = wave_range_sum(P.W.Pa, 0, P.nb);
    // S at the chunk's first entry: v'_0 = -tail plus the entries >= 1 before the chunk
    const double start = blk == 0 ? 0.0 : -tail + wave_range_sum(P.W.Pa, 0, blk);
    double e[MPER], tot = 0.0;
#pragma unroll
    for (int k = 0; k < MPER; k++) {
        const int i = base + k;
        e[k] = (i == 0) ? -tail : (i < N ? P.in[i] : 0.0);
        tot += e[k];
    }
    double run = start + mb_offset(tot, lds);
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < MPER; k++) {
        const int i = base + k;
        if (i < N) {
            P.W.S[i] = run;
            ssum += run;
        }
        run += e[k];
    }
    const double t2 = mb_sum(ssum, lds);
    if (threadIdx.x == 0) P.W.Pb[blk] = t2;
}

__global__ __launch_bounds__(MT) void k_pm_scan_y(PoissonMB P) {
    __shared__ double lds[MW];
    const int N = P.fs.n_f, blk = blockIdx.x, base = blk * CHUNK + threadIdx.x * MPER;
    const double d0 = wave_range_sum(P.W.Pb, 0, P.nb) * P.invN;
    // sum of D over the earlier chunks: (entries - [chunk 0]) * d0 - (chunk sum of S)
    double sd = 0.0;
    for (int k = threadIdx.x & 63; k < blk; k += 64) {
        const int n = min(CHUNK, N - k * CHUNK) - (k == 0 ? 1 : 0);
        sd += (double)n * d0 - P.W.Pb[k];
    }
    const double offD = wave_sum(sd);
    double D[MPER], tot = 0.0;
#pragma unroll
    for (int k = 0; k < MPER; k++) {
        const int i = base + k;
        D[k] = (i == 0 || i >= N) ? 0.0 : d0 - P.W.S[i];
        tot += D[k];
    }
    double run = offD + mb_offset(tot, lds);
#pragma unroll
    for (int k = 0; k < MPER; k++) {
        const int i = base + k;
        run += D[k];
        if (i < N) P.W.y[i] = run;
    }
}

// MODE 0: first solve (x = c); 1: refinement (x += c); both then form the next
// residual.  MODE 2: last refinement, writes phi (x_out), Ex0 and E.
template <int MODE>
__global__ __launch_bounds__(MT) void k_pm_conv(PoissonMB P) {
    __shared__ double yw[CHUNK + 2 * CW];
    __shared__ double ys[NSPEC + 2 * (NG - 1)];
    __shared__ double xw[CHUNK + 2 * HALO];
    __shared__ double cs[NSPEC];
    __shared__ double lds[MW];
    const vx_field_state &F = P.fs;
    const int N = F.n_f, blk = blockIdx.x, lo = blk * CHUNK;
    const int n_own = min(CHUNK, N - lo);
    const double *y = P.W.y;
    for (int j = threadIdx.x; j < n_own + 2 * CW; j += MT) yw[j] = y[wrap(lo - CW + j, N)];
    for (int j = threadIdx.x; j < NSPEC + 2 * (NG - 1); j += MT) ys[j] = y[wrap(-3 - (NG - 1) + j, N)];
    __syncthreads();
    if (threadIdx.x < NSPEC) cs[threadIdx.x] = conv_at(ys, threadIdx.x + NG - 1, P.g);
    __syncthreads();
    const double cp0 = cs[3];  // (M^-1 y)_0
    // row 0 (the approximate solve's last step): c_0 from c_1, c_2, c_{N-2}, c_{N-1}
    const double c0 = P.in[0] - (((P.w1 * (cs[4] - cp0) + P.w2 * (cs[5] - cp0)) + P.w2 * (cs[1] - cp0)) +
                                 P.w1 * (cs[2] - cp0));
    // x on the chunk and its halo
    for (int j = threadIdx.x; j < n_own + 2 * HALO; j += MT) {
        const int gi = wrap(lo - HALO + j, N);
        const double c = (gi == 0) ? c0 : conv_at(yw, j + NG - 1, P.g) - cp0;
        const double xn = MODE == 0 ? c : P.x_in[gi] + c;
        xw[j] = xn;
        if (j >= HALO && j < n_own + HALO) P.x_out[gi] = xn;
    }
    __syncthreads();
    if (MODE < 2) {
        // r = b - A x in double-double, rounded once (term order as the oracle's)
        double s = 0.0;
        for (int j = threadIdx.x; j < n_own; j += MT) {
            const int i = lo + j, c = j + HALO;
            double sh = P.W.b[i], sl = 0.0;
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
                sub(1.0, xw[c]); sub(P.w1, xw[c + 1]); sub(P.w2, xw[c + 2]); sub(P.w2, xw[c - 2]);
                sub(P.w1, xw[c - 1]);
            } else {
                const double cf[5] = {P.w2, P.w1, P.w3, P.w1, P.w2};
                for (int d = -2; d <= 2; d++) {
                    if (wrap(i + d, N) == 0) continue;
                    sub(cf[d + 2], xw[c + d]);
                }
            }
            const double r = sh + sl;
            P.r_out[i] = r;
            if (i >= 1) s += r;
        }
        const double tot = mb_sum(s, lds);
        if (threadIdx.x == 0) P.W.Pa[blk] = tot;
    } else {
        // EMSolver.cpp:191: Ex0 += -(E(-1) + E(0)) * 0.5 with the entry Ex0, from phi
        // at N-3..N-1, 0..2 (recomputed here with the owners' operations)
        double ps[NSPEC];
#pragma unroll
        for (int k = 0; k < NSPEC; k++) {
            const int gi = wrap(k - 3, N);
            ps[k] = P.x_in[gi] + (k == 3 ? c0 : cs[k] - cp0);
        }
        const double ex0 = *P.W.ex0, fc = F.field_coef;
        const double e_m1 = -fc * (8 * (ps[3] - ps[1]) - ps[4] + ps[0]) + ex0;
        const double e_0 = -fc * (8 * (ps[4] - ps[2]) - ps[5] + ps[1]) + ex0;
        const double ex0n = ex0 + -(e_m1 + e_0) * 0.5;
        if (blk == 0 && threadIdx.x == 0) *F.Ex0 = ex0n;
        for (int j = threadIdx.x; j < n_own; j += MT) {
            const int c = j + HALO;
            F.E[lo + j] = -fc * (8 * (xw[c + 1] - xw[c - 1]) - xw[c + 2] + xw[c - 2]) + ex0n;
        }
    }
}

// Register/LDS variant for n_f <= 4 * PT: thread t owns the CH consecutive
// entries [t*CH, t*CH + CH); neighbours (residual, convolution) come from LDS.
//
// Barriers: a workgroup sum or scan exchanges the 16 wavefront partials through
// LDS once (one barrier); every thread then forms the same sequential sums of
// the partials itself.  Consecutive exchanges alternate between two partial
// buffers, so a buffer is rewritten only after a later barrier has ordered
// every read of it.
constexpr int SMALL_MAX = 4 * PT;

struct SmallLds {
    double v[SMALL_MAX];  // neighbour exchange: y for the convolution, x for the residual
    double part[2][PW];   // wavefront partials, alternating
    double edge[5];       // c at entries 0, 1, 2, N-2, N-1 (row 0 of the matrix)
};

// Workgroup exclusive scan of the per-thread totals (thread order), and the
// workgroup total; every thread gets the same total bits.
__device__ __forceinline__ double block_scan(double tot, double (&part)[2][PW], int &buf, double *total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const double inc = wave_incl_scan(tot);
    if (lane == 63) part[buf][w] = inc;
    __syncthreads();
    double off = 0.0, all = 0.0;
#pragma unroll
    for (int k = 0; k < PW; k++) {
        const double v = part[buf][k];
        if (k == w) off = all;
        all += v;
    }
    buf ^= 1;
    if (total) *total = all;
    return off + (inc - tot);
}

__device__ __forceinline__ double block_sum(double v, double (&part)[2][PW], int &buf) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    v = wave_sum(v);
    if (lane == 0) part[buf][w] = v;
    __syncthreads();
    double all = 0.0;
#pragma unroll
    for (int k = 0; k < PW; k++) all += part[buf][k];
    buf ^= 1;
    return all;
}

// One approximate solve of A out = in (the steps of the multi-workgroup
// variant above): S = exclusive prefix of (-sum_{k>=1} in_k, in_1, in_2, ...),
// D = mean(S) - S (D_0 = 0), y = inclusive prefix of D, c = M^-1 y - (M^-1 y)_0,
// and c_0 from row 0 of the matrix.
template <int CH>
__device__ void approx_small(const PoissonArgs &P, const double *in, double *out, SmallLds &L, int &buf) {
    const int N = P.fs.n_f;
    const int base = threadIdx.x * CH;
    double loc = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k >= 1 && base + k < N) loc += in[k];
    const double tail = block_sum(loc, L.part, buf);
    double rp[CH], S[CH];
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        rp[k] = (i == 0) ? -tail : (i < N ? in[k] : 0.0);
        tot += rp[k];
    }
    double run = block_scan(tot, L.part, buf, nullptr);
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        S[k] = run;
        run += rp[k];
        if (base + k < N) ssum += S[k];
    }
    const double d0 = block_sum(ssum, L.part, buf) * P.invN;
    double D[CH];
    tot = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        int i = base + k;
        D[k] = (i == 0 || i >= N) ? 0.0 : d0 - S[k];
        tot += D[k];
    }
    run = block_scan(tot, L.part, buf, nullptr);
#pragma unroll
    for (int k = 0; k < CH; k++) {
        run += D[k];
        if (base + k < N) L.v[base + k] = run;
    }
    __syncthreads();
    // the thread's CH entries share one window of CH + 32 taps:
    // w[q] = y[base - 16 + q], periodic (below 34 entries the taps can wrap
    // more than once)
    constexpr int WN = CH + 2 * (NG - 1);
    double cp[CH];
    if (base < N) {
        double w[WN];
        if (N >= 2 * NG) {
#pragma unroll
            for (int q = 0; q < WN; q++) {
                int u = base - (NG - 1) + q;
                u = u < 0 ? u + N : (u >= N ? u - N : u);
                w[q] = L.v[u];
            }
        } else {
#pragma unroll
            for (int q = 0; q < WN; q++) {
                int u = base - (NG - 1) + q;
                while (u >= N) u -= N;
                while (u < 0) u += N;
                w[q] = L.v[u];
            }
        }
#pragma unroll
        for (int k = 0; k < CH; k++) {
            double v = P.g[0] * w[NG - 1 + k];
#pragma unroll
            for (int m = 1; m < NG; m++) v += P.g[m] * (w[NG - 1 + k + m] + w[NG - 1 + k - m]);
            cp[k] = base + k < N ? v : 0.0;
        }
    } else {
#pragma unroll
        for (int k = 0; k < CH; k++) cp[k] = 0.0;
    }
    // row 0 needs c at entries 0, 1, 2, N-2 and N-1
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = base + k;
        if (i == 0) L.edge[0] = cp[k];
        if (i == 1) L.edge[1] = cp[k];
        if (i == 2) L.edge[2] = cp[k];
        if (i == N - 2) L.edge[3] = cp[k];
        if (i == N - 1) L.edge[4] = cp[k];
    }
    __syncthreads();
    const double c0 = L.edge[0];
#pragma unroll
    for (int k = 0; k < CH; k++) out[k] = (base + k > 0) ? cp[k] - c0 : cp[k];
    if (threadIdx.x == 0) {
        const double y1 = L.edge[1] - c0, y2 = L.edge[2] - c0, yn2 = L.edge[3] - c0, yn1 = L.edge[4] - c0;
        out[0] = in[0] - (((P.w1 * y1 + P.w2 * y2) + P.w2 * yn2) + P.w1 * yn1);
    }
}

// r = b - A x with a double-double accumulation of the reference's
// coefficients (rows 1..N-1: the periodic stencil without column 0; row 0:
// x_0 + w1 x_1 + w2 x_2 + w2 x_{N-2} + w1 x_{N-1})
template <int CH>
__device__ void residual_small(const PoissonArgs &P, const double *b, const double *x, double *r, SmallLds &L) {
    const int N = P.fs.n_f;
    const int base = threadIdx.x * CH;
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k < N) L.v[base + k] = x[k];
    __syncthreads();
    const double *xb = L.v;
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
}

template <int CH>
__global__ __launch_bounds__(PT) void k_poisson_small(PoissonArgs P) {
    const vx_field_state &F = P.fs;
    const int N = F.n_f;
    const int t = threadIdx.x, base = t * CH;
    __shared__ SmallLds L;
    int buf = 0;
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
    approx_small<CH>(P, b, x, L, buf);
    for (int it = 0; it < 3; it++) {
        // the previous solve's last barrier ordered its reads of L.v
        residual_small<CH>(P, b, x, r, L);
        // the first barrier of the solve below orders the residual's reads of
        // L.v before the solve writes it
        approx_small<CH>(P, r, c, L, buf);
#pragma unroll
        for (int k = 0; k < CH; k++) x[k] = x[k] + c[k];
    }
    // E from the final potential: every thread's reads of L.v in the last
    // solve came before its final barrier
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k < N) {
            L.v[base + k] = x[k];
            F.phi[base + k] = x[k];
        }
    __syncthreads();
    const double *xb = L.v;
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
    const double *drive;  // when set, by0 = drive[0] and bz0 = drive[1] in device memory
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
        const double by0 = A.drive ? A.drive[0] : A.by0, bz0 = A.drive ? A.drive[1] : A.bz0;
        rEy = -dx_inv * (Bz[i1] - bz0) * vx_eps0_inv * vx_mu_inv;
        rEz = dx_inv * (By[i1] - by0) * vx_eps0_inv * vx_mu_inv;
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

// ------------------------------------------------ fixed work array of vx_poisson_noalloc
// The multi-workgroup solve's work vector (pm_work layout) for n_f up to
// VX_POISSON_NOALLOC_MAX_NF: 3.67 MB that each GPU the library runs on holds
// until the process ends.  Host code never uses this name as an address: the
// device address comes from hipGetSymbolAddress, once per device, with that
// device current (scratch_address, at context creation).
constexpr size_t NOALLOC_WORK =
    7 * (size_t)VX_POISSON_NOALLOC_MAX_NF + 2 * (size_t)pm_chunks(VX_POISSON_NOALLOC_MAX_NF) + 1;
static __device__ __align__(256) double scratch[NOALLOC_WORK];
constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mu;
static double *scratch_ptr[MAX_DEVICES] = {};

int validate_fs(vx_ctx *ctx, const vx_field_state *fs) {
    if (!fs) return fail(ctx, VX_EINVAL, "null field state");
    if (fs->n_f < 8 || fs->n_prepad < 2 || fs->n_postpad < 2 ||
        fs->n_e != fs->n_f + fs->n_prepad + fs->n_postpad)
        return fail(ctx, VX_EINVAL, "bad field sizes n_f=%d pre=%d post=%d n_e=%d", fs->n_f,
                    fs->n_prepad, fs->n_postpad, fs->n_e);
    return VX_OK;
}

}  // namespace

// Called by vx_ctx_create with ctx->device current: the device address of
// this GPU's fixed work array (looked up once per device), or a recorded error.
int vx_noalloc_scratch_init(vx_ctx *ctx) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES) return VX_EINVAL;
    if (!scratch_ptr[ctx->device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return VX_EHIP;
        }
        scratch_ptr[ctx->device] = (double *)p;
    }
    ctx->poisson_scratch = scratch_ptr[ctx->device];
    return VX_OK;
}

extern "C" {

size_t vx_field_work_elems(int32_t n_f) { return (size_t)7 * n_f + 2 * (size_t)pm_chunks(n_f) + 1; }

static int poisson(vx_ctx *ctx, const vx_field_state *fs, double *work) {
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
        return check_launch(ctx, "k_poisson");
    }
    PoissonMB M;
    M.fs = *fs;
    M.w1 = P.w1;
    M.w2 = P.w2;
    M.w3 = P.w3;
    for (int k = 0; k < NG; k++) M.g[k] = P.g[k];
    M.invN = P.invN;
    M.nb = pm_chunks(fs->n_f);
    if (!work) return fail(ctx, VX_EINVAL, "poisson: no work array for n_f = %d", fs->n_f);
    M.W = pm_work(work, fs->n_f);
    const PmWork &W = M.W;
    const unsigned nb = (unsigned)M.nb;
    k_pm_rhs<<<nb, MT, 0, ctx->stream>>>(M);
    // four approximate solves: of b (x = c), then of three residuals (x += c)
    const double *in[4] = {W.b, W.ra, W.rb, W.ra};
    double *r_out[4] = {W.ra, W.rb, W.ra, nullptr};
    const double *x_in[4] = {nullptr, W.xa, W.xb, W.xa};
    double *x_out[4] = {W.xa, W.xb, W.xa, fs->phi};
    for (int s = 0; s < 4; s++) {
        M.in = in[s];
        M.r_out = r_out[s];
        M.x_in = x_in[s];
        M.x_out = x_out[s];
        k_pm_scan_s<<<nb, MT, 0, ctx->stream>>>(M);
        k_pm_scan_y<<<nb, MT, 0, ctx->stream>>>(M);
        if (s == 0) k_pm_conv<0><<<nb, MT, 0, ctx->stream>>>(M);
        else if (s < 3) k_pm_conv<1><<<nb, MT, 0, ctx->stream>>>(M);
        else k_pm_conv<2><<<nb, MT, 0, ctx->stream>>>(M);
    }
    return check_launch(ctx, "k_poisson");
}


int vx_poisson(vx_ctx *ctx, const vx_field_state *fs) { return poisson(ctx, fs, fs ? fs->work : nullptr); }

int vx_poisson_noalloc(vx_ctx *ctx, const vx_field_state *fs) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    if (fs->n_f > VX_POISSON_NOALLOC_MAX_NF)
        return fail(ctx, VX_EINVAL, "vx_poisson_noalloc: n_f = %d exceeds its fixed work array (n_f <= %d)", fs->n_f,
                    VX_POISSON_NOALLOC_MAX_NF);
    if (!ctx->poisson_scratch) return fail(ctx, VX_EHIP, "vx_poisson_noalloc: no work array address on device %d", ctx->device);
    return poisson(ctx, fs, ctx->poisson_scratch);
}

}  // extern "C"

static int maxwell_stage(vx_ctx *ctx, const vx_field_state *fs, int step, double dt, double by0,
                         double bz0, const double *drive) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    if (step < 0 || step > 5) return fail(ctx, VX_EINVAL, "maxwell: bad step %d", step);
    MaxArgs A;
    A.fs = *fs;
    A.step = step;
    A.by0 = by0;
    A.bz0 = bz0;
    A.drive = drive;
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

extern "C" {

int vx_maxwell_stage(vx_ctx *ctx, const vx_field_state *fs, int step, double dt, double by0,
                     double bz0) {
    return maxwell_stage(ctx, fs, step, dt, by0, bz0, nullptr);
}

int vx_maxwell_stage_dev(vx_ctx *ctx, const vx_field_state *fs, int step, double dt,
                         const double *drive) {
    if (!drive) return fail(ctx, VX_EINVAL, "maxwell: null drive");
    return maxwell_stage(ctx, fs, step, dt, 0.0, 0.0, drive);
}

int vx_neutralize(vx_ctx *ctx, const vx_field_state *fs) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    k_neutralize<<<(unsigned)((fs->n_f + 255) / 256), 256, 0, ctx->stream>>>(*fs);
    return check_launch(ctx, "k_neutralize");
}

}  // extern "C"

// EMSolver.cpp:631-664 on the device: pc = max_i (the per-cell term), each term
// with the host code's operations; max is exact, so the order does not matter.
constexpr int CFL_MAXSP = 8;
struct CflArgs {
    vx_field_state fs;
    int n_species;
    double aq[CFL_MAXSP], m[CFL_MAXSP], dps[CFL_MAXSP];  // |q_j|, m_j, 1 / dp_j
    double dpsMax;
};

__global__ __launch_bounds__(1024) void k_cfl(CflArgs A, double *out) {
    const vx_field_state &F = A.fs;
    __shared__ double part[16];
    double pc = 0.0;
    for (int i = threadIdx.x; i < F.n_f; i += 1024) {
        const double Azv = F.Az[i], Ayv = F.Ay[i];
        const double As = Ayv * Ayv + Azv * Azv;
        double temp = 0.0;
        for (int j = 0; j < A.n_species; j++) {
            const double v = A.aq[j] / A.m[j] / sqrt(1 + As / ((A.m[j] * vx_cs) * (A.m[j] * vx_cs))) * A.dps[j];
            temp = (temp < v) ? v : temp;
        }
        const double t = temp * fabs(Ayv * F.Bz[i] - Azv * F.By[i]) + A.dpsMax * fabs(F.E[i]);
        pc = (pc < t) ? t : pc;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(pc, off, 64);
        pc = (pc < o) ? o : pc;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = pc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int k = 0; k < 16; k++) r = (r < part[k]) ? part[k] : r;
        *out = r;
    }
}

extern "C" {

int vx_cfl_bound_async(vx_ctx *ctx, const vx_field_state *fs, int n_species, const double *q, const double *m,
                       const double *dp_finest) {
    int rc = validate_fs(ctx, fs);
    if (rc) return rc;
    if (n_species < 1 || n_species > CFL_MAXSP) return fail(ctx, VX_EINVAL, "cfl: %d species", n_species);
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "cfl bound while a step graph is recorded");
    if (!ctx->cfl_dev) {
        if ((rc = check_hip(ctx, hipMalloc((void **)&ctx->cfl_dev, sizeof(double)), "cfl buffer"))) return rc;
        if ((rc = check_hip(ctx, hipHostMalloc((void **)&ctx->cfl_host, sizeof(double)), "cfl pinned"))) return rc;
        if ((rc = check_hip(ctx, hipEventCreateWithFlags(&ctx->cfl_ev, hipEventDisableTiming), "cfl event")))
            return rc;
    }
    CflArgs A;
    A.fs = *fs;
    A.n_species = n_species;
    A.dpsMax = 0.0;
    for (int i = 0; i < n_species; i++) {  // EMSolver.cpp:637-642
        A.aq[i] = std::fabs(q[i]);
        A.m[i] = m[i];
        A.dps[i] = 1 / dp_finest[i];
        A.dpsMax = std::max(A.dpsMax, std::fabs(q[i]) * A.dps[i]);
    }
    k_cfl<<<1, 1024, 0, ctx->stream>>>(A, ctx->cfl_dev);
    if ((rc = check_launch(ctx, "k_cfl"))) return rc;
    rc = check_hip(ctx, hipMemcpyAsync(ctx->cfl_host, ctx->cfl_dev, sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream), "cfl copy");
    if (rc) return rc;
    if ((rc = check_hip(ctx, hipEventRecord(ctx->cfl_ev, ctx->stream), "cfl event record"))) return rc;
    ctx->cfl_c_dx = vx_cs / fs->dx_f;
    ctx->cfl_pending = true;
    return VX_OK;
}

int vx_cfl_ready(vx_ctx *ctx, int *ready) {
    if (!ctx->cfl_pending || !ready) return fail(ctx, VX_EINVAL, "cfl: no bound pending");
    hipError_t e = hipEventQuery(ctx->cfl_ev);
    if (e == hipErrorNotReady) {
        *ready = 0;
        return VX_OK;
    }
    *ready = 1;
    return check_hip(ctx, e, "cfl query");
}

int vx_cfl_result(vx_ctx *ctx, double *out) {
    if (!ctx->cfl_pending || !out) return fail(ctx, VX_EINVAL, "cfl: no bound pending");
    int rc = check_hip(ctx, hipEventSynchronize(ctx->cfl_ev), "cfl wait");
    if (rc) return rc;
    ctx->cfl_pending = false;
    *out = 1.0 / std::max((ctx->cfl_c_dx + *ctx->cfl_host), 1e-40);
    return VX_OK;
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
