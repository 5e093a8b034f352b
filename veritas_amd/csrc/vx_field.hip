// 1-D field solver on the finest grid (EMFieldSolver, EMSolver.cpp).
//
// Poisson (EMSolver.cpp:28-88,156-192).  The reference factors the dense N x N
// matrix with LAPACK and solves it at every RK stage, O(N^2) per solve.  Here
// the same matrix is solved to near correct rounding in O(N) (one 512-thread
// workgroup with the vector in registers up to N = 4096, a sequence of
// multi-workgroup launches above):
//   rows 1..N-1 of the matrix are the periodic 4th-order stencil
//   (w2, w1, w3, w1, w2) with column 0 removed, i.e. the periodic operator with
//   phi_0 pinned to 0; row 0 then gives phi_0.  The periodic operator factors
//   exactly as L4 = L2 * M with L2 = [-1, 2, -1] and M = [-1, 14, -1] / 12, so
//   an approximate solve is two prefix sums (L2^-1 on zero-mean data) and M^-1,
//   whose entries decay as (7 - 4 sqrt 3)^|k| (a 33-tap convolution, or two
//   first-order recurrences).
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

constexpr int NG = 17;    // convolution half-width + 1 for M^-1

struct PoissonArgs {
    vx_field_state fs;
    double w1, w2, w3;
    double g[NG];  // M^-1 coefficients, g[|k|]
    double rho, C;  // M^-1 = C (1 - rho S^-1)^-1 (1 - rho S)^-1
    double invN;
};


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

// Single-workgroup variant for n_f <= SMALL_MAX: thread t owns the CH
// consecutive entries [t*CH, t*CH + CH) in registers; neighbours come from an
// LDS copy of the vector with a periodic halo of VH entries on either side, so
// no stencil or recurrence needs index wrapping.
//
// M^-1 = C (1 - rho S^-1)^-1 (1 - rho S)^-1 with C = 12 rho / (1 - rho^2), so
// (M^-1 y)_i = C (u_i + v_i - y_i) with the first-order recurrences
// u_i = y_i + rho u_{i-1} and v_i = y_i + rho v_{i+1}.  Each thread runs them
// over its entries after VH warm-up entries (rho^17 < 4e-20, the reach of the
// convolution the multi-workgroup variant uses).  The approximate solve only
// has to be accurate: the refinement rounds take x to the correctly rounded
// solution whatever the rounding inside it.
//
// Barriers: a workgroup sum or scan exchanges the PW1 wavefront partials
// through LDS once (one barrier); every thread then forms the same sums of the
// partials itself.  Consecutive exchanges alternate between two partial
// buffers, so a buffer is rewritten only after a later barrier has ordered
// every read of it.
constexpr int PT1 = 512;             // threads of the single-workgroup solve
constexpr int PW1 = PT1 / 64;        // its wavefronts
constexpr int SMALL_MAX = 8 * PT1;   // largest n_f it serves (CH = 8)
constexpr int VH = NG - 1;           // periodic halo of the LDS copy

constexpr int VLEN = SMALL_MAX + 2 * VH + 8;  // positions of the LDS copy
constexpr int VDUMMY = VLEN - VH;             // at(VDUMMY): a slot past them, written only

struct SmallLds {
    // at(j) = vec[j mod N] for j in [-VH, N + VH); the 8 positions after that
    // are 0 (the last thread's backward recurrence may start there).  Position
    // p = VH + j is stored at p + p / 8: a thread reads 8 consecutive entries
    // per step of its loops, and the pad after every 8 puts the lanes of a
    // wavefront (9 doubles apart) in different LDS banks.
    double v[(VLEN + 1) + (VLEN + 1) / 8 + 1];
    double part[2][PW1];  // wavefront partials, alternating
    double edge[5];       // c at entries 0, 1, 2, N-2, N-1 (row 0 of the matrix)
    __device__ __forceinline__ double &at(int j) {
        const int p = VH + j;
        return v[p + (p >> 3)];
    }
};

// DPP move of a double; lanes outside row_mask or without a source lane get 0
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_or_0(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Inclusive scan over the wavefront's lanes: within rows of 16 by row shifts
// of 1, 2, 4, 8, then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), then rows 0-1's total into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ double wave_scan_dpp(double v) {
    v += dpp_or_0<0x111>(v);
    v += dpp_or_0<0x112>(v);
    v += dpp_or_0<0x114>(v);
    v += dpp_or_0<0x118>(v);
    v += dpp_or_0<0x142, 0xa>(v);
    v += dpp_or_0<0x143, 0xc>(v);
    return v;
}

__device__ __forceinline__ double lane63(double v) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Workgroup sum of v; every thread gets the same bits.
__device__ __forceinline__ double block_sum1(double v, double (&part)[2][PW1], int &buf) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double tot = lane63(wave_scan_dpp(v));
    if (lane == 0) part[buf][w] = tot;
    __syncthreads();
    double all = 0.0;
#pragma unroll
    for (int k = 0; k < PW1; k++) all += part[buf][k];
    buf ^= 1;
    return all;
}

// Workgroup exclusive scan of the per-thread totals in thread order.
__device__ __forceinline__ double block_excl_scan1(double tot, double (&part)[2][PW1], int &buf) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double inc = wave_scan_dpp(tot);
    if (lane == 63) part[buf][w] = inc;
    const double exc = dpp_or_0<0x138>(inc);  // wave_shr:1: lane l - 1's inclusive sum, 0 at lane 0
    __syncthreads();
    double off = 0.0;
#pragma unroll
    for (int k = 0; k < PW1; k++)
        if (k < w) off += part[buf][k];
    buf ^= 1;
    return off + exc;
}

// The thread's entries into the LDS copy, with their periodic images in the
// halo.  For N >= 2 VH an entry has at most one image (i + N for i < VH,
// i - N for i >= N - VH), stored without a branch (an entry without an image
// stores itself twice); below that (N >= 8) the images are looped over, a
// branch the whole workgroup takes alike.
template <int CH>
__device__ __forceinline__ void put_periodic(SmallLds &L, const double *val, int N, int base) {
    if (N >= 2 * VH) {
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int i = base + k;
            if (i < N) {
                const int q = i < VH ? i + N : (i >= N - VH ? i - N : i);
                L.at(i) = val[k];
                L.at(q) = val[k];
            }
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = base + k;
        if (i >= N) continue;
        L.at(i) = val[k];
        for (int q = i - N; q >= -VH; q -= N) L.at(q) = val[k];
        for (int q = i + N; q < N + VH; q += N) L.at(q) = val[k];
    }
}

// One approximate solve of A out = in: S = exclusive prefix of
// (-sum_{k>=1} in_k, in_1, in_2, ...), D = mean(S) - S (D_0 = 0), y = inclusive
// prefix of D, c = M^-1 y - (M^-1 y)_0, and c_0 from row 0 of the matrix.
template <int CH>
__device__ void approx_one(const PoissonArgs &P, const double *in, double *out, SmallLds &L, int &buf) {
    const int N = P.fs.n_f;
    const int base = threadIdx.x * CH;
    double loc = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (base + k >= 1 && base + k < N) loc += in[k];
    const double tail = block_sum1(loc, L.part, buf);
    double rp[CH], S[CH];
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = base + k;
        rp[k] = (i == 0) ? -tail : (i < N ? in[k] : 0.0);
        tot += rp[k];
    }
    double run = block_excl_scan1(tot, L.part, buf);
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        S[k] = run;
        run += rp[k];
        if (base + k < N) ssum += S[k];
    }
    const double d0 = block_sum1(ssum, L.part, buf) * P.invN;
    double D[CH];
    tot = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = base + k;
        D[k] = (i == 0 || i >= N) ? 0.0 : d0 - S[k];
        tot += D[k];
    }
    run = block_excl_scan1(tot, L.part, buf);
    double y[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        run += D[k];
        y[k] = run;
    }
    // (every thread's reads of L.v before this solve came before its first
    // barrier above)
    put_periodic<CH>(L, y, N, base);
    __syncthreads();
    double cp[CH];
    if (base < N) {
        const double rho = P.rho;
        double fw[CH], bw[CH];
        double u = L.at(base - VH);
#pragma unroll
        for (int m = 1; m < VH; m++) u = fma(rho, u, L.at(base - VH + m));
#pragma unroll
        for (int k = 0; k < CH; k++) {
            u = fma(rho, u, L.at(base + k));
            fw[k] = u;
        }
        double b = L.at(base + CH - 1 + VH);
#pragma unroll
        for (int m = VH - 1; m >= 1; m--) b = fma(rho, b, L.at(base + CH - 1 + m));
#pragma unroll
        for (int k = CH - 1; k >= 0; k--) {
            b = fma(rho, b, L.at(base + k));
            bw[k] = b;
        }
#pragma unroll
        for (int k = 0; k < CH; k++) cp[k] = base + k < N ? P.C * ((fw[k] + bw[k]) - y[k]) : 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < CH; k++) cp[k] = 0.0;
    }
    // row 0 needs c at entries 0, 1, 2, N-2 and N-1 (N >= 8: distinct)
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = base + k;
        const int e = i < 3 ? i : (i == N - 2 ? 3 : (i == N - 1 ? 4 : -1));
        if (e >= 0) L.edge[e] = cp[k];
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
__device__ void residual_one(const PoissonArgs &P, const double *b, const double *x, double *r, SmallLds &L) {
    const int N = P.fs.n_f;
    const int base = threadIdx.x * CH;
    put_periodic<CH>(L, x, N, base);
    __syncthreads();
    // x at entries base - 2 .. base + CH + 1, read at once (inside the LDS
    // array for every thread; entries >= N are not used)
    double w[CH + 4];
#pragma unroll
    for (int m = 0; m < CH + 4; m++) w[m] = L.at(base - 2 + m);
    auto sub = [](double &sh, double &sl, double a, double xq) {
        const double p = a * xq;
        const double e = fma(a, xq, -p);
        const double tt = sh - p;
        const double bv = tt - sh;
        const double err = (sh - (tt - bv)) + (-p - bv);
        sl = sl + (err - e);
        sh = tt;
    };
    const double cf[5] = {P.w2, P.w1, P.w3, P.w1, P.w2};
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = base + k;
        // rows i >= 1: neighbours j = i + d in [-1, N + 1]; column 0 (j = 0 or
        // j = N) gets a zero coefficient, which changes neither sum
        double sh = b[k], sl = 0.0;
#pragma unroll
        for (int d = -2; d <= 2; d++) {
            const int j = i + d;
            sub(sh, sl, (j == 0 || j == N) ? 0.0 : cf[d + 2], w[k + d + 2]);
        }
        r[k] = i < N ? sh + sl : 0.0;
    }
    if (base == 0) {  // row 0
        double sh = b[0], sl = 0.0;
        sub(sh, sl, 1.0, L.at(0));
        sub(sh, sl, P.w1, L.at(1));
        sub(sh, sl, P.w2, L.at(2));
        sub(sh, sl, P.w2, L.at(N - 2));
        sub(sh, sl, P.w1, L.at(N - 1));
        r[0] = sh + sl;
    }
}

template <int CH>
__global__ __launch_bounds__(PT1) void k_poisson_one(PoissonArgs P) {
    const vx_field_state &F = P.fs;
    const int N = F.n_f;
    const int t = threadIdx.x, base = t * CH;
    __shared__ SmallLds L;
    int buf = 0;
    // the entries past the last image, where the last thread's backward
    // recurrence may start, hold 0 (never written again)
    if (t < 8) L.at(N + VH + t) = 0.0;
    double b[CH], x[CH], r[CH], c[CH];
    const double te = vx_eps0_inv;
    // the right-hand side formed at entries t + PT1 m (coalesced loads, at
    // clamped indices, all in flight at once) and handed to the owners of
    // consecutive entries through the LDS copy
    double chv[CH], nev[CH];
#pragma unroll
    for (int m = 0; m < CH; m++) {
        const int i = min(t + PT1 * m, N - 1);
        chv[m] = F.charge[i];
        nev[m] = F.neutral[i];
    }
    // the entry Ex0 (written back at the end) as a vector load behind them,
    // so that its latency overlaps theirs (a scalar load would be waited for
    // before them)
    const double ex0 = __builtin_bit_cast(
        double, __builtin_amdgcn_raw_buffer_load_b64(
                    __builtin_amdgcn_make_buffer_rsrc(F.Ex0, 0, 8, 0x00020000), 0, 0, 0));
#pragma unroll
    for (int m = 0; m < CH; m++) {
        const int i = t + PT1 * m;
        double v = te * (chv[m] + nev[m]);
        v *= F.dx2;
        L.at(i < N ? i : VDUMMY) = v;  // no branch around the loads' use
    }
    __syncthreads();
    // (these reads come before the first barrier of the first solve, which
    // writes the LDS copy only after it)
#pragma unroll
    for (int k = 0; k < CH; k++) {
        b[k] = base + k < N ? L.at(base + k) : 0.0;
        x[k] = 0.0;
    }
    // the approximate solve of b, then three rounds of refinement (residual r
    // of x, x += the approximate solve of r), as one loop: the kernel runs
    // once per field solve and its code is fetched cold, so one copy of each
    // step is fetched once and then reused
#pragma nounroll
    for (int it = 0; it < 4; it++) {
        if (it > 0) {
            // the previous solve's last barrier ordered its reads of L.v
            residual_one<CH>(P, b, x, r, L);
        } else {
#pragma unroll
            for (int k = 0; k < CH; k++) r[k] = b[k];
        }
        // the first barrier of the solve below orders the residual's reads of
        // L.v before the solve writes it
        approx_one<CH>(P, r, c, L, buf);
#pragma unroll
        for (int k = 0; k < CH; k++) x[k] = it == 0 ? c[k] : x[k] + c[k];
    }
    // phi and E from the final potential (EMSolver.cpp:137-154, 191), stored
    // at entries t + PT1 m (coalesced) from the LDS copy: every thread's reads
    // of it in the last solve came before its final barrier
    put_periodic<CH>(L, x, N, base);
    __syncthreads();
    const double fc = F.field_coef;
    auto ef = [&](double e0, double pm2, double pm1, double pp1, double pp2) {
        return -fc * (8 * (pp1 - pm1) - pp2 + pm2) + e0;
    };
    const double e_m1 = ef(ex0, L.at(-3), L.at(-2), L.at(0), L.at(1)), e_0 = ef(ex0, L.at(-2), L.at(-1), L.at(1), L.at(2));
    const double ex0n = ex0 + -(e_m1 + e_0) * 0.5;
    if (t == 0) *F.Ex0 = ex0n;
#pragma unroll
    for (int m = 0; m < CH; m++) {
        const int i = t + PT1 * m;
        if (i < N) {
            F.phi[i] = L.at(i);
            F.E[i] = ef(ex0n, L.at(i - 2), L.at(i - 1), L.at(i + 1), L.at(i + 2));
        }
    }
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

// EMSolver.cpp:204-338; one thread per element and field (blockIdx.y): the
// six fields' updates are independent
__global__ void k_maxwell_update(MaxArgs A) {
    const vx_field_state &F = A.fs;
    const int ne = F.n_e;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    const int v = blockIdx.y;
    double *y = v == 0 ? F.By : v == 1 ? F.Bz : v == 2 ? F.Ey : v == 3 ? F.Ez : v == 4 ? F.Ay : F.Az;
    if (A.step <= 4) {
        double t = y[fidx(ne, i, 0)] + A.a[0] * y[fidx(ne, i, 2)];
        for (int k = 1; k <= A.step; k++) t = t + A.a[k] * y[fidx(ne, i, k + 2)];
        y[fidx(ne, i, 1)] = t;
    } else {
        double t = y[fidx(ne, i, 1)] + A.b[0] * y[fidx(ne, i, 2)];
        for (int k = 1; k < 6; k++) t = t + A.b[k] * y[fidx(ne, i, k + 2)];
        y[fidx(ne, i, 0)] = t;
        y[fidx(ne, i, 1)] = t;
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
    P.rho = rho;
    P.C = C;
    P.invN = 1.0 / fs->n_f;
    LaunchTimer lt(ctx, "k_poisson");
    const int ch = (fs->n_f + PT1 - 1) / PT1;
    if (fs->n_f <= SMALL_MAX) {
        if (ch <= 1) k_poisson_one<1><<<1, PT1, 0, ctx->stream>>>(P);
        else if (ch <= 2) k_poisson_one<2><<<1, PT1, 0, ctx->stream>>>(P);
        else if (ch <= 4) k_poisson_one<4><<<1, PT1, 0, ctx->stream>>>(P);
        else k_poisson_one<8><<<1, PT1, 0, ctx->stream>>>(P);
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
    k_maxwell_update<<<dim3(g, 6), 256, 0, ctx->stream>>>(A);
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
