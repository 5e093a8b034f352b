// p-moments: charge and current density (Rectangle::CalculateRhoAndJ,
// Rectangle.cpp:157-282) and their assembly on the finest grid
// (Level.cpp:19-62, Mesh.cpp:52-56, EMSolver.cpp:104-122).
//
// The reference sums over p sequentially (j = 0, 1, ..., n_p-1) for each x
// column and each finest sub-cell k.  To reproduce those bits one wavefront
// owns one column: its 64 lanes evaluate the summands of 64 consecutive p-cells
// in parallel (interpolants, gamma factors, logarithms) and stage them in LDS;
// lane 0 then adds them in order.  4096 columns give 4096 independent chains,
// which keeps every CU busy while each chain stays in reference order.
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <vector>

#include "vx_dev.h"
#include "vx_internal.h"
#include "vx_math.h"

using namespace vxi;
using namespace vxd;

namespace {

constexpr int MAX_RTB = 4;  // coarsest level of a 3-level hierarchy at ratio 2
constexpr int TAB_KEY = 8;  // A^2, pmin, dp, p_pos, mc2inv, c1, c2, n_p

// Rectangle.cpp:76-89
__constant__ double kInterp[12] = {
    0.104166666666667, -0.708333333333334, 0.708333333333334, -0.104166666666667,
    0.117647058823529, 0.029411764705882,  0.029411764705882, 0.117647058823529,
    -0.083333333333333, 0.166666666666667, -0.166666666666667, 0.083333333333334};

struct MomArgs {
    vx_patch p;
    vx_fields fl;
    double c1, c2, c3;     // m_inv*c_inv, 1/c1, 1/48
    double inv_rtb;       // 1.0 / relativeToBottom
    double dpq, jscale;   // dp*q and -q*q/mass
    double rel[3 * MAX_RTB];  // interpolatCoefsREL (Rectangle.cpp:103-110)
    // Two tables of logarithm terms, each for one value a0 of A^2: table 0 for
    // the patch's first sub-column (every column when A^2 is uniform), table 1
    // for A^2 = 0 (the plasma ahead of a laser pulse).  Table q starts at
    // tab + q (n_p + 3 + TAB_KEY): entries t in [0, n_p+2) hold
    // c2 log(h(t)/h(t-1)) with h(j) = Gamma(p_j, a0) + c1 p_j, entry n_p+2 is 1
    // once built, and the TAB_KEY entries after it what they were formed from
    // (a0 first, then the p grid and the constants), so an unchanged table is
    // not rebuilt.  A wavefront whose chains all have A^2 == a0 of a table
    // reads it (the same bits it would form itself); the others form their
    // logarithms per distinct A^2.
    const double *tab;
};

// EMFieldSolver::GetCellAverageASquared (EMSolver.hpp:57-64)
__device__ __forceinline__ double cell_a2(const vx_fields &fl, long i) {
    i += fl.n_prepad;
    i = i > -1 ? i : 0;
    i = i < fl.n_e ? i : (fl.n_e - 1);
    double ay = fl.Ay1[i], az = fl.Az1[i];
    return (ay * ay) + (az * az);
}


__host__ __device__ inline long tab_stride(int n_p) { return (long)n_p + 3 + TAB_KEY; }

// Logarithm tables of one patch (see MomArgs::tab); workgroup q builds table q.
template <int RTB>
__global__ __launch_bounds__(1024) void k_moment_table(MomArgs A, double *tab) {
    const vx_patch &P = A.p;
    const int t = threadIdx.x;
    tab += blockIdx.x * tab_stride(P.n_p);
    const double a0 = blockIdx.x == 0 ? P.q * P.q * cell_a2(A.fl, (long)P.x_pos * RTB) : 0.0;
    const double key[TAB_KEY] = {a0, P.pmin, P.dp, (double)P.p_pos, P.mc2inv, A.c1, A.c2, (double)P.n_p};
    double *kt = tab + P.n_p + 3;
    bool hit = tab[P.n_p + 2] == 1.0;
    for (int k = 0; k < TAB_KEY; k++) hit = hit && kt[k] == key[k];
    if (__syncthreads_or(hit ? 1 : 0) != 0) return;  // the same table is in place
    // (every thread read the same bits; the barrier also orders those reads
    // before the writes below)
    auto h = [&](int j) {
        double pm = momentum(P.pmin, P.dp, j, P.p_pos);
        return gamma_(pm, a0, P.mc2inv) + A.c1 * pm;
    };
    for (int e = t; e < P.n_p + 2; e += blockDim.x) tab[e] = A.c2 * vx_log(h(e) / h(e - 1));
    if (t < TAB_KEY) kt[t] = key[t];
    if (t == 0) tab[P.n_p + 2] = 1.0;
}

// LDS exchange inside one wavefront: order the lanes' LDS writes before the
// other lanes' reads (no workgroup barrier is needed; waves are independent).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Compensated wavefront sum of the lanes' (value, error) pairs (mixed-precision
// form, where the reference's p order is not kept): a butterfly of error-free
// additions (TwoSum); every lane ends with the same sum s and error e.
__device__ __forceinline__ void wave_two_sum(double v, double ve, double &s, double &e) {
    s = v;
    e = ve;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double so = __shfl_xor(s, off, 64), eo = __shfl_xor(e, off, 64);
        const double t = s + so, z = t - s;
        const double err = (s - (t - z)) + (so - z);
        e = (e + eo) + err;
        s = t;
    }
}

// (acc, cmp) += v with an error-free addition
__device__ __forceinline__ void add_two(double &acc, double &cmp, double v) {
    const double t = acc + v, z = t - acc;
    const double err = (acc - (t - z)) + (v - z);
    cmp = cmp + err;
    acc = t;
}

// A wavefront owns C x columns (NCH = C * RTB finest sub-columns, the
// reference's accumulation chains) and walks p in chunks of 64.  Per chunk the
// 64 lanes (one per p-cell) evaluate the summands of every chain in parallel
// and stage them in LDS; then lane ch (< NCH) adds the 64 charge summands of
// chain ch in p order and lane 32 + ch the current summands.  Skipped cells
// (nested, or past n_p) contribute -0.0, which is an exact identity for
// addition, so every chain performs the reference's additions in its order.
//
// Latency: everything a chunk's summands read from global memory (the cell
// values, the logarithm table, the nested flags) is loaded two chunks ahead
// into registers and staged in LDS, so the summands and the ordered sums read
// LDS only and never wait on the next chunks' loads.
//
// FUSE (plain form only): the ordered sums of chunk k run in the same
// straight-line block as the summands of chunk k + 1 (two summand buffers),
// so the latency of each dependent addition is filled with the next chunk's
// arithmetic; this holds many values in registers, and is the choice for
// launches of at most one wavefront per SIMD.  Otherwise the summands and the
// ordered sums of a chunk run one after the other through one buffer, in
// fewer registers and less LDS, so that more wavefronts share each SIMD.
//
// PAIR (plain form only): two wavefronts per column group, a producer that
// stages chunk k and writes its summands while the consumer adds chunk k - 1
// in order, one workgroup barrier per chunk.  The SIMD interleaves the two
// instruction streams, which hides the latency of each (for launches of at
// most one group per SIMD).
template <int RTB, int C, bool COMP, bool FUSE, bool PAIR = false>
__device__ __forceinline__ void moments_body(MomArgs A) {
    constexpr int NCH = C * RTB;
    static_assert(NCH <= 32, "chains must fit in half a wavefront");
    const vx_patch &P = A.p;
    static_assert(!PAIR || (!COMP && !FUSE), "the pair form is a plain form of its own");
    constexpr int NW = PAIR ? 1 : 4;  // column groups per workgroup
    const int lane = threadIdx.x & 63;
    // LDS slot of this wavefront's column group; in the pair form wavefront 0
    // produces and wavefront 1 consumes the group's summands
    const int wave = PAIR ? 0 : (int)(threadIdx.x >> 6);
    const bool producer = !PAIR || threadIdx.x < 64, consumer = !PAIR || threadIdx.x >= 64;
    const int i0 = (PAIR ? (int)blockIdx.x : (int)blockIdx.x * 4 + wave) * C;
    __shared__ double S[NW][NCH][68];   // sub-cell values, p index jb-1+e
    __shared__ double H[NW][68];        // Gamma(p)+c1 p for the chain being prepared
    __shared__ double Lb[NW][NCH][68];  // c2 log(h(j+1)/h(j)), p index jb-1+e
    __shared__ double T[FUSE || PAIR ? 2 : 1][NW][2][NCH][65];  // summands: buffer, group, charge/current, chain, p
    if (i0 >= P.n_x) return;  // uniform per wavefront (per workgroup in the pair form)
    const int pitch = P.pitch;
    const double *g = P.f[1];
    double asq[NCH];
    for (int ch = 0; ch < NCH; ch++) {
        int c = ch / RTB, k = ch % RTB;
        int i = i0 + c < P.n_x ? i0 + c : P.n_x - 1;
        asq[ch] = P.q * P.q * cell_a2(A.fl, (long)(i + P.x_pos) * RTB + k);
    }
    // every chain of this wavefront has the A^2 of one of the tables: tabulated logarithms
    const double *tab = nullptr;
    for (int q = 1; q >= 0 && !tab; q--) {
        const double *tq = A.tab + q * tab_stride(P.n_p);
        bool all = tq[P.n_p + 2] != 0.0;
        for (int ch = 0; ch < NCH; ch++) all = all && asq[ch] == tq[P.n_p + 3];
        if (all) tab = tq;
    }
    const bool use_tab = tab != nullptr;
    double acc = 0.0;  // lanes < NCH: charge chain; lanes 32..32+NCH-1: current chain
    // COMP: per chain, this lane's running sum over its p-cells and its compensation
    double accR[NCH], cmpR[NCH], accJ[NCH], cmpJ[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) accR[ch] = cmpR[ch] = accJ[ch] = cmpJ[ch] = 0.0;
    // sub-cell values of the RTB chains of column c at p-cell j (Rectangle.cpp:
    // 183-204, 258-274): GetInterpolantsREL(f1..f5)[k] (Rectangle.cpp:139-155)
    // plus the correction that makes the sub-cell values average to the cell
    // value (Rectangle.cpp:195-204), formed once per cell for all k
    // (i0 + c, j) is a cell whose value the sums use (others count as 0)
    auto cell_ok = [&](int c, int j) { return i0 + c < P.n_x && j < P.n_p + 2; };
    auto column = [&](int c, int j, double *out) {
        // loads at indices clamped into the plane, with no branches, so that
        // the loads of the chunks in flight are counted one by one and a chunk
        // waits only for its own
        const int i = i0 + c;
        const bool ok = cell_ok(c, j);
        const int ic = i < P.n_x ? i : P.n_x - 1, jc = j < P.n_p + 2 ? j : P.n_p + 1;
        if (RTB == 1) {
            // interpolant coefficients and correction are exactly 0 for
            // relativeToBottom == 1: the sub-cell value is the cell value,
            // masked by cell_ok where it is staged (not here, which would wait
            // for the load)
            out[0] = g[pidx(pitch, ic, jc)];
            return;
        }
        double f1 = g[pidx(pitch, ic - 2, jc)], f2 = g[pidx(pitch, ic - 1, jc)];
        const double f3 = g[pidx(pitch, ic, jc)];
        double f4 = g[pidx(pitch, ic + 1, jc)], f5 = g[pidx(pitch, ic + 2, jc)];
        f5 -= f3; f4 -= f3; f2 -= f3; f1 -= f3;
        const double a1 = kInterp[0] * f1 + kInterp[1] * f2 + kInterp[2] * f4 + kInterp[3] * f5;
        const double a2 = kInterp[4] * f1 + kInterp[5] * f2 + kInterp[6] * f4 + kInterp[7] * f5;
        const double a3 = kInterp[8] * f1 + kInterp[9] * f2 + kInterp[10] * f4 + kInterp[11] * f5;
        double v[RTB], te = 0.0;
#pragma unroll
        for (int k = 0; k < RTB; k++) {
            v[k] = A.rel[3 * k] * a1 + A.rel[3 * k + 1] * a2 + A.rel[3 * k + 2] * a3 + f3;
            te += v[k] + 0.0;  // the interpolant with a zero correction
        }
        const double corr = f3 - A.inv_rtb * te;
#pragma unroll
        for (int k = 0; k < RTB; k++) out[k] = ok ? v[k] + corr : 0.0;
    };
    // a chunk in registers: values for p in [j0, j0+66), the table entries
    // tab[j0+1 .. j0+67) (clamped to the table) and the nested flag of each
    // column's p-cell j0+1+lane; two such sets are in flight (two chunks ahead)
    struct Pre {
        double v[NCH][2], t[2];
        int n[C];
    };
    auto prefetch = [&](int j0, Pre &pr) {
        for (int c = 0; c < C; c++) {
            double v[RTB];
            column(c, j0 + lane, v);
            for (int k = 0; k < RTB; k++) pr.v[c * RTB + k][0] = v[k];
            column(c, j0 + 64 + min(lane, 1), v);  // read from lanes 0 and 1 only
            for (int k = 0; k < RTB; k++) pr.v[c * RTB + k][1] = v[k];
        }
        // without a table or nested flags the loads read the plane instead
        // (and are not used): no branch around them
        const double *tq = use_tab ? tab : g;
        const int t0 = min(j0 + 1 + lane, P.n_p + 2), t1 = min(j0 + 65 + min(lane, 1), P.n_p + 2);
        pr.t[0] = tq[t0];
        pr.t[1] = tq[t1];  // read from lanes 0 and 1 only
        const uint8_t *nq = P.nested ? P.nested : (const uint8_t *)g;
        const int jn = min(j0 + 1 + lane, P.n_p - 1);
        for (int c = 0; c < C; c++) pr.n[c] = nq[pidx(pitch, min(i0 + c, P.n_x - 1), jn)];
    };
    // value set `set` of chain ch held for p-cell j (0 outside the sums' cells)
    auto staged = [&](const Pre &pr, int ch, int j, int set) {
        return (RTB > 1 || cell_ok(ch / RTB, j)) ? pr.v[ch][set] : 0.0;
    };
    // summands of p-cell j = jb + lane for every chain (Rectangle.cpp:214-253),
    // into T[b] (or the lanes' compensated sums); no branches and LDS reads only
    auto summands = [&](int jb, int b, const int *nflag) {
        const int j = jb + lane;
        const bool live = j < P.n_p;
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            const int c = ch / RTB;
            const bool skip = !live || i0 + c >= P.n_x || nflag[c] != 0;
            const double *L = use_tab ? &Lb[wave][0][0] : &Lb[wave][ch][0];
            const double tm = S[wave][ch][lane], t0 = S[wave][ch][lane + 1], tp = S[wave][ch][lane + 2];
            const double gm = L[lane + 1], gm1 = L[lane], gp1 = L[lane + 2];
            const double jv = t0 * gm + A.c3 * (gp1 - gm1) * (tp - tm);
            const double tr = skip ? -0.0 : t0, tj = skip ? -0.0 : jv;
            if (COMP) {
                add_two(accR[ch], cmpR[ch], tr);
                add_two(accJ[ch], cmpJ[ch], tj);
            } else {
                T[b][wave][0][ch][lane] = tr;
                T[b][wave][1][ch][lane] = tj;
            }
        }
    };
    // entries e + 1 and e + 2 of a 66-entry sequence whose entries 0..63 are v0
    // of lanes 0..63 and entries 64, 65 are v1 of lanes 0 and 1 (lane e's view)
    auto next2 = [&](double v0, double v1, double &n1, double &n2) {
        const double x64 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v1), 0),
                                            __builtin_amdgcn_readlane(__double2loint(v1), 0));
        const double x65 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v1), 1),
                                            __builtin_amdgcn_readlane(__double2loint(v1), 1));
        const double s1 = shdn1(v0), s2 = shdn1(s1);
        n1 = lane == 63 ? x64 : s1;
        n2 = lane == 63 ? x65 : (lane == 62 ? x64 : s2);
    };
    // summands as above from registers (pair form with a logarithm table): the
    // same operands, so the same bits, without staging them in LDS
    auto summands_reg = [&](int jb, int b, const int *nflag, const double (*sv)[2], double st0, double st1) {
        const int j = jb + lane;
        const bool live = j < P.n_p;
        double gm, gp1;
        next2(st0, st1, gm, gp1);
        const double gm1 = st0;
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            const int c = ch / RTB;
            const bool skip = !live || i0 + c >= P.n_x || nflag[c] != 0;
            double t0, tp;
            next2(sv[ch][0], sv[ch][1], t0, tp);
            const double tm = sv[ch][0];
            const double jv = t0 * gm + A.c3 * (gp1 - gm1) * (tp - tm);
            T[b][wave][0][ch][lane] = skip ? -0.0 : t0;
            T[b][wave][1][ch][lane] = skip ? -0.0 : jv;
        }
    };
    // FUSE: the ordered sums of buffer b ^ 1 (the previous chunk) with
    // the summands of chunk jb (into buffer b) placed between their groups of
    // eight additions, so that the additions' latency is filled with them
    auto fused = [&](int jb, int b, const int *nflag) {
        const int j = jb + lane;
        const bool live = j < P.n_p;
        const int chr = (lane & 31) < NCH ? (lane & 31) : 0;
        const double *row = &T[b ^ 1][wave][lane >> 5][chr][0];
        double a = acc, v[16];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = row[q];
#pragma unroll
        for (int g = 0; g < 8; g++) {
            double *cur = &v[(g & 1) * 8], *nxt = &v[(~g & 1) * 8];
            if (g + 1 < 8) {
#pragma unroll
                for (int q = 0; q < 8; q++) nxt[q] = row[8 * (g + 1) + q];
            }
#pragma unroll
            for (int ch = g * NCH / 8; ch < (g + 1) * NCH / 8; ch++) {
                const int c = ch / RTB;
                const bool skip = !live || i0 + c >= P.n_x || nflag[c] != 0;
                const double *L = use_tab ? &Lb[wave][0][0] : &Lb[wave][ch][0];
                const double tm = S[wave][ch][lane], t0 = S[wave][ch][lane + 1], tp = S[wave][ch][lane + 2];
                const double gm = L[lane + 1], gm1 = L[lane], gp1 = L[lane + 2];
                const double jv = t0 * gm + A.c3 * (gp1 - gm1) * (tp - tm);
                T[b][wave][0][ch][lane] = skip ? -0.0 : t0;
                T[b][wave][1][ch][lane] = skip ? -0.0 : jv;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) a += cur[q];
        }
        acc = a;
    };
    // chain (lane & 31) of charge (lanes < 32) or current adds the 64 summands
    // of T[b] in p order; the other lanes repeat chain 0 and are not stored
    auto ordered = [&](int b) {
        const int ch = (lane & 31) < NCH ? (lane & 31) : 0;
        const double *row = &T[b][wave][lane >> 5][ch][0];
        double a = acc;
#pragma unroll
        for (int e = 0; e < 64; e += 8) {
            double v0 = row[e], v1 = row[e + 1], v2 = row[e + 2], v3 = row[e + 3];
            double v4 = row[e + 4], v5 = row[e + 5], v6 = row[e + 6], v7 = row[e + 7];
            a += v0; a += v1; a += v2; a += v3; a += v4; a += v5; a += v6; a += v7;
        }
        acc = a;
    };
    if (!COMP && FUSE) {
        // the ordered sums of the (absent) chunk before the first add -0.0
        for (int e = lane; e < 2 * NCH * 65; e += 64) (&T[FUSE ? 1 : 0][wave][0][0][0])[e] = -0.0;
    }
    // Logarithm terms of chain ch for chunk jb: Lb[wave][ch][e] = L(jb + e),
    // e in [0, 66), with L(m) = c2 log(h(m) / h(m-1)) and h(m) = Gamma(p_m) + c1 p_m
    // for the chain's A^2.  The first chunk forms h at all 67 p-cells it needs;
    // each later one takes L(jb), L(jb + 1) from the previous chunk's entries 64
    // and 65 and h(jb + 1) from hcar, so that its 64 lanes form one h and one
    // logarithm each (h(m - 1) from the next lower lane): the same operands, so
    // the same bits, for half the arithmetic.
    double hcar[NCH];  // h(jb + 65) of the last chunk, per chain (uniform)
    auto log_row = [&](int jb, int ch) {
        double *Lrow = &Lb[wave][ch][0];
        if (jb == 0) {
            for (int e = lane; e < 67; e += 64) {
                double pm = momentum(P.pmin, P.dp, jb - 1 + e, P.p_pos);
                H[wave][e] = gamma_(pm, asq[ch], P.mc2inv) + A.c1 * pm;
            }
            wave_sync();
            for (int e = lane; e < 66; e += 64) Lrow[e] = A.c2 * vx_log(H[wave][e + 1] / H[wave][e]);
            hcar[ch] = H[wave][66];
            wave_sync();
            return;
        }
        const double pm = momentum(P.pmin, P.dp, jb + 2 + lane, P.p_pos);
        const double hm = gamma_(pm, asq[ch], P.mc2inv) + A.c1 * pm;
        // the lane shift runs with every lane active (a DPP move from a lane
        // outside the mask reads 0), then lane 0 takes the carried h
        const double hs = shup1(hm);
        const double hp = lane == 0 ? hcar[ch] : hs;
        const double lv = A.c2 * vx_log(hm / hp);
        const double keep = Lrow[64 + min(lane, 1)];  // L(jb), L(jb + 1) for lanes 0, 1
        wave_sync();
        if (lane < 2) Lrow[lane] = keep;
        Lrow[2 + lane] = lv;
        hcar[ch] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(hm), 63),
                                    __builtin_amdgcn_readlane(__double2loint(hm), 63));
        wave_sync();
    };
    // chunk jb from pr (summand buffer bb), whose registers then take chunk jb + 128
    auto chunk = [&](int jb, int bb, Pre &pr) {
        for (int ch = 0; ch < NCH; ch++) {
            S[wave][ch][lane] = staged(pr, ch, jb - 1 + lane, 0);
            if (lane < 2) S[wave][ch][64 + lane] = staged(pr, ch, jb + 63 + lane, 1);
        }
        if (use_tab) {
            Lb[wave][0][lane] = pr.t[0];
            if (lane < 2) Lb[wave][0][64 + lane] = pr.t[1];
        }
        int nflag[C];
        for (int c = 0; c < C; c++) nflag[c] = P.nested ? pr.n[c] : 0;
        if (jb + 128 < P.n_p) prefetch(jb + 127, pr);
        // logarithm terms, computed once per distinct A^2 among the chains
        for (int ch = 0; ch < NCH && !use_tab; ch++) {
            if (ch > 0 && asq[ch] == asq[ch - 1]) {
                for (int e = lane; e < 66; e += 64) Lb[wave][ch][e] = Lb[wave][ch - 1][e];
                continue;
            }
            log_row(jb, ch);
        }
        wave_sync();
        if (COMP) {
            summands(jb, bb, nflag);
        } else if (FUSE) {
            fused(jb, bb, nflag);
        } else {
            summands(jb, 0, nflag);
            wave_sync();
            ordered(0);
        }
        wave_sync();
    };
    Pre pa, pb;
    if (producer) {
        prefetch(-1, pa);
        prefetch(63, pb);  // clamped loads: also when there is no second chunk
    }
    int last = 0;  // summand buffer of the last chunk
    // pair form: the producer fills buffer bb with chunk jb while the consumer
    // adds the other buffer (chunk jb - 64); the barrier hands them over.  Each
    // wavefront runs a loop of its own (the same number of barriers), and the
    // producer's loop is specialised on the table case, so that no branch joins
    // the registers of the chunks in flight (a join would copy them, and the
    // copy would wait for the loads).
    auto prod_chunk = [&](auto tab_c, int jb, int bb, Pre &pr) __attribute__((always_inline)) {
        constexpr bool TAB = decltype(tab_c)::value;
        if constexpr (TAB) {
            // the chunk's cell values and table entries stay in registers: the
            // neighbours p + 1 and p + 2 come from DPP lane shifts, and the two
            // entries past the last lane from lanes 0 and 1 of the second set
            double sv[NCH][2];
            for (int ch = 0; ch < NCH; ch++) {
                sv[ch][0] = staged(pr, ch, jb - 1 + lane, 0);
                sv[ch][1] = staged(pr, ch, jb + 63 + min(lane, 1), 1);
            }
            const double st0 = pr.t[0], st1 = pr.t[1];
            int nflag[C];
            for (int c = 0; c < C; c++) nflag[c] = P.nested ? pr.n[c] : 0;
            prefetch(jb + 127, pr);  // clamped loads: also past the last chunk
            summands_reg(jb, bb, nflag, sv, st0, st1);
        } else {
            for (int ch = 0; ch < NCH; ch++) {
                S[0][ch][lane] = staged(pr, ch, jb - 1 + lane, 0);
                if (lane < 2) S[0][ch][64 + lane] = staged(pr, ch, jb + 63 + lane, 1);
            }
            int nflag[C];
            for (int c = 0; c < C; c++) nflag[c] = P.nested ? pr.n[c] : 0;
            prefetch(jb + 127, pr);  // clamped loads: also past the last chunk
            for (int ch = 0; ch < NCH; ch++) {
                if (ch > 0 && asq[ch] == asq[ch - 1]) {
                    for (int e = lane; e < 66; e += 64) Lb[0][ch][e] = Lb[0][ch - 1][e];
                    continue;
                }
                log_row(jb, ch);  // wave is 0 in the pair form
            }
            wave_sync();
            summands(jb, bb, nflag);
        }
        __syncthreads();
    };
    auto prod_loop = [&](auto tab_c) __attribute__((always_inline)) {
        for (int jb = 0; jb < P.n_p; jb += 128) {
            prod_chunk(tab_c, jb, 0, pa);
            if (jb + 64 < P.n_p) prod_chunk(tab_c, jb + 64, 1, pb);
        }
    };
    if (PAIR && producer) {
        if (use_tab) prod_loop(std::true_type{});
        else prod_loop(std::false_type{});
    } else if (PAIR) {
        for (int jb = 0; jb < P.n_p; jb += 128) {
            if (jb > 0) ordered(1);
            __syncthreads();
            last = 0;
            if (jb + 64 < P.n_p) {
                ordered(0);
                __syncthreads();
                last = 1;
            }
        }
    } else {
        for (int jb = 0; jb < P.n_p; jb += 128) {
            chunk(jb, 0, pa);
            last = 0;
            if (jb + 64 < P.n_p) {
                chunk(jb + 64, 1, pb);
                last = 1;
            }
        }
    }
    if (!COMP && (FUSE || PAIR) && consumer) ordered(last);
    if (COMP) {
        // one compensated reduction across the lanes per chain
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            double s, e;
            wave_two_sum(accR[ch], cmpR[ch], s, e);
            accR[ch] = s;
            cmpR[ch] = e;
            wave_two_sum(accJ[ch], cmpJ[ch], s, e);
            accJ[ch] = s;
            cmpJ[ch] = e;
        }
    }
    if (consumer && (lane & 31) < NCH) {
        int ch = lane & 31;
        int c = ch / RTB, k = ch % RTB;
        int i = i0 + c;
        if (i < P.n_x) {
            double sum = acc;
            if (COMP) {
#pragma unroll
                for (int q = 0; q < NCH; q++)
                    if (q == ch) sum = lane < 32 ? accR[q] + cmpR[q] : accJ[q] + cmpJ[q];
            }
            if (lane < 32) P.chargeR[(long)i * RTB + k] = sum * A.dpq;
            else P.currentR[(long)i * RTB + k] = sum * A.jscale;
        }
    }
}

template <int RTB, int C, bool COMP, bool FUSE, bool PAIR = false>
__global__ __launch_bounds__(256) void k_moments(MomArgs A) {
    moments_body<RTB, C, COMP, FUSE, PAIR>(A);
}
// The same kernel compiled for four resident waves per SIMD (at most 128
// VGPRs): the split form of relativeToBottom-4 patches fits without spilling
// (137 VGPRs, three waves, otherwise); the other forms would spill.
template <int RTB, int C, bool COMP, bool FUSE, bool PAIR = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_moments_w4(MomArgs A) {
    moments_body<RTB, C, COMP, FUSE, PAIR>(A);
}

// Wide form for relativeToBottom 1 (one chain per column and moment): a
// workgroup owns WCOL consecutive x-columns.  Producer wavefront c stages the
// summands of column c for one 64-cell p-chunk at a time in LDS (the same
// operands as k_moments: the cell values of p-cells j-1, j, j+1 and the
// logarithm terms of j, j+1, j+2 from the next lanes by DPP shifts), with the
// loads of the next WDEPTH chunks in flight; the consumer wavefront adds the
// previous chunk in p order, lane c the charge chain of column c and lane
// 32 + c its current chain, all 2 WCOL chains in lockstep, so one wavefront's
// dependent additions serve WCOL columns.  A workgroup barrier per chunk hands
// the buffers over.  The same bits as k_moments<1, C>, for patches without
// nested cells (the finest level: its cells are never covered by a finer one).
constexpr int WCOL = 8;

// Whole-wavefront rotation by one lane (DPP wave_rol:1): lane l receives the
// value of lane (l + 1) mod 64.
__device__ __forceinline__ double rol1(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Entries l + 1 and l + 2 (lane l's view) of the sequence cur[0..63],
// nxt[0], nxt[1] (this chunk's values, then the next chunk's first two).
__device__ __forceinline__ void next2_rot(int lane, double cur, double nxt, double &n1, double &n2) {
    n1 = rol1(lane == 0 ? nxt : cur);
    n2 = rol1(rol1(lane < 2 ? nxt : cur));
}

// Buffer resource over nbytes bytes from p (raw: no stride, no swizzle); a
// load past nbytes returns 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *p, int nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, nbytes, 0x00020000);
}
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// COMP (mixed-precision form): no consumer; each producer keeps its lanes'
// compensated (TwoSum) sums over the chunks and ends with one compensated
// reduction across its lanes, as the compensated k_moments form does.
template <int WDEPTH, bool COMP>
__global__ __launch_bounds__(64 * (WCOL + 1)) void k_moments_wide(MomArgs A) {
    const vx_patch &P = A.p;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = (int)(threadIdx.x & 63);
    const int i0 = (int)blockIdx.x * WCOL;
    // buffer, charge/current, column, p-cell of the chunk (rows of 65: the
    // consumer's lanes read one row each, in different LDS banks)
    __shared__ double T[2][2][WCOL][65];
    // chunks k < nchunk hold p-cells; the loops run NK iterations, a multiple
    // of WDEPTH, so that the producers' unrolled loop has no exit inside
    // (chunks past nchunk consist of skipped cells: -0.0 summands, which
    // change no sum)
    const int nchunk = (P.n_p + 63) / 64;
    const int NK = (nchunk + 1 + WDEPTH - 1) / WDEPTH * WDEPTH;
    if (wave < WCOL) {
        const int i = i0 + wave;
        const bool col_ok = i < P.n_x;
        const int ic = col_ok ? i : P.n_x - 1;
        // this column's A^2 and a logarithm table built for it, if there is one
        const double asq = P.q * P.q * cell_a2(A.fl, (long)(ic + P.x_pos));
        const double *tab = nullptr;
        for (int q = 1; q >= 0 && !tab; q--) {
            const double *tq = A.tab + q * tab_stride(P.n_p);
            if (tq[P.n_p + 2] != 0.0 && asq == tq[P.n_p + 3]) tab = tq;
        }
        // Chunk k (jb = 64 k) loads entry r = jb + l of two sequences: the
        // column's f from p = -1 (f at p-cell jb - 1 + l) and the logarithm
        // table (entry jb + l), with r clamped to n_p + 2 (f at p-cell n_p + 1,
        // the table's last entry); those clamped entries serve skipped cells
        // only.  The two values past the chunk's last lane (f at jb + 63,
        // jb + 64; table entries jb + 64, jb + 65) are lanes 0 and 1 of chunk
        // k + 1.  Lane offsets are min(l, n_p + 2 - jb'), jb' = min(jb, n_p + 2),
        // and jb' is the scalar offset.
        const __amdgpu_buffer_rsrc_t rf = raw_rsrc(P.f[1] + pidx(P.pitch, ic, -1), 8 * (P.n_p + 3));
        const __amdgpu_buffer_rsrc_t rt = raw_rsrc(tab ? tab : P.f[1], 8 * (P.n_p + 3));
        auto h = [&](int m) {
            const double pm = momentum(P.pmin, P.dp, m, P.p_pos);
            return gamma_(pm, asq, P.mc2inv) + A.c1 * pm;
        };
        // the column's own logarithm terms c2 log(h(m)/h(m-1)) at m = jb + l
        auto logs = [&](int jb) { return A.c2 * vx_log(h(jb + lane) / h(jb + lane - 1)); };
        // The loop is specialised on the table case, so that no branch joins
        // the registers of the chunks in flight (a join makes the loads wait),
        // and its ring of chunks in flight is unrolled: slot d holds chunk
        // k0 + d in fixed registers and is reloaded with chunk k0 + d + WDEPTH
        // once used (a rotating copy would wait for the loads in flight).
        double aR = 0.0, cR = 0.0, aJ = 0.0, cJ = 0.0;  // COMP: the lanes' sums
        auto run = [&](auto tab_c) __attribute__((always_inline)) {
            constexpr bool TAB = decltype(tab_c)::value;
            double pv[WDEPTH], pt[WDEPTH];
            auto load = [&](int k, int d) __attribute__((always_inline)) {
                const int jl = min(64 * k, P.n_p + 2);
                const int vo = 8 * min(lane, P.n_p + 2 - jl), so = 8 * jl;
                pv[d] = buf_load(rf, vo, so);
                if (TAB) pt[d] = buf_load(rt, vo, so);
            };
            // the slots are loaded in the order they are used (the scheduling
            // barriers keep the compiler from reordering them), so that waiting
            // for slot d leaves the later slots' loads in flight
#pragma unroll
            for (int d = 0; d < WDEPTH; d++) {
                load(d, d);
                __builtin_amdgcn_sched_barrier(0);
            }
            double lc = TAB ? 0.0 : logs(0);  // no table: chunk k's logarithm terms
            for (int k0 = 0; k0 < NK; k0 += WDEPTH) {
#pragma unroll
                for (int d = 0; d < WDEPTH; d++) {
                    // chunk k (every summand skipped at k >= nchunk): no branch,
                    // so the wait for each slot's loads counts only the slots
                    // loaded after it
                    const int k = k0 + d, dn = (d + 1) % WDEPTH;  // dn: chunk k + 1
                    const int jb = 64 * k;
                    double gm1, gnx;
                    if (TAB) {
                        gm1 = pt[d];
                        gnx = pt[dn];
                    } else {
                        gm1 = lc;
                        gnx = logs(jb + 64);
                        lc = gnx;
                    }
                    double t0, tp, gm, gp1;
                    next2_rot(lane, pv[d], pv[dn], t0, tp);
                    next2_rot(lane, gm1, gnx, gm, gp1);
                    const double tm = pv[d];
                    const bool skip = (jb + lane >= P.n_p) | !col_ok;
                    const double jv = t0 * gm + A.c3 * (gp1 - gm1) * (tp - tm);
                    if (COMP) {
                        add_two(aR, cR, skip ? -0.0 : t0);
                        add_two(aJ, cJ, skip ? -0.0 : jv);
                    } else {
                        T[k & 1][0][wave][lane] = skip ? -0.0 : t0;
                        T[k & 1][1][wave][lane] = skip ? -0.0 : jv;
                    }
                    load(k + WDEPTH, d);
                    __builtin_amdgcn_sched_barrier(0);
                    if (!COMP) __syncthreads();
                }
            }
        };
        if (tab) run(std::true_type{});
        else run(std::false_type{});
        if (COMP) {
            double sr, er, sj, ej;
            wave_two_sum(aR, cR, sr, er);
            wave_two_sum(aJ, cJ, sj, ej);
            if (lane == 0 && col_ok) {
                P.chargeR[i] = (sr + er) * A.dpq;
                P.currentR[i] = (sj + ej) * A.jscale;
            }
        }
    } else if (!COMP) {
        // consumer: chain (lane & 31) of charge (lanes < 32) or current; lanes
        // past WCOL repeat chain 0 and are not stored
        // the dependent additions are the critical path of each chunk: the
        // consumer issues ahead of the producers that share its SIMD
        __builtin_amdgcn_s_setprio(3);
        const int ch = (lane & 31) < WCOL ? (lane & 31) : 0;
        const int kind = lane >> 5;
        double acc = 0.0;
        for (int k = 0; k < NK; k++) {
            if (k > 0) {
                const double *row = &T[(k - 1) & 1][kind][ch][0];
                double a = acc;
#pragma unroll
                for (int e = 0; e < 64; e += 8) {
                    const double v0 = row[e], v1 = row[e + 1], v2 = row[e + 2], v3 = row[e + 3];
                    const double v4 = row[e + 4], v5 = row[e + 5], v6 = row[e + 6], v7 = row[e + 7];
                    a += v0; a += v1; a += v2; a += v3; a += v4; a += v5; a += v6; a += v7;
                }
                acc = a;
            }
            __syncthreads();
        }
        const int c = lane & 31;
        if (c < WCOL && i0 + c < P.n_x) {
            if (lane < 32) P.chargeR[i0 + c] = acc * A.dpq;
            else P.currentR[i0 + c] = acc * A.jscale;
        }
    }
}

// Rectangle::CalculateEnergy (Rectangle.cpp:284-305) without its data race:
// chain (j, k) sums the p-sub-cell value k of every non-nested cell (i, j) in
// increasing i, then scales by dx (the reference's OpenMP loop adds to the same
// energyR entries from several threads; the oracle and this kernel use the
// sequential order).
template <int RTB>
__global__ __launch_bounds__(256) void k_energy(MomArgs A, double *energyR) {
    const vx_patch &P = A.p;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.n_p * RTB) return;
    const int j = e / RTB, k = e % RTB;
    const int pitch = P.pitch;
    const double *g = P.f[1];
    double acc = 0.0;
    for (int i = 0; i < P.n_x; i++) {
        if (P.nested && P.nested[pidx(pitch, i, j)]) continue;
        double f1 = g[pidx(pitch, i, j - 2)], f2 = g[pidx(pitch, i, j - 1)], f3 = g[pidx(pitch, i, j)];
        double f4 = g[pidx(pitch, i, j + 1)], f5 = g[pidx(pitch, i, j + 2)];
        f5 -= f3; f4 -= f3; f2 -= f3; f1 -= f3;
        double a1 = kInterp[0] * f1 + kInterp[1] * f2 + kInterp[2] * f4 + kInterp[3] * f5;
        double a2 = kInterp[4] * f1 + kInterp[5] * f2 + kInterp[6] * f4 + kInterp[7] * f5;
        double a3 = kInterp[8] * f1 + kInterp[9] * f2 + kInterp[10] * f4 + kInterp[11] * f5;
        acc += A.rel[3 * k] * a1 + A.rel[3 * k + 1] * a2 + A.rel[3 * k + 2] * a3 + f3;
    }
    energyR[(long)j * RTB + k] = acc * P.dx;
}

struct MomentSrc {
    const double *chargeR, *currentR;
    int start, count;  // finest-grid span [start, start + count)
};

// Assembly in the reference order: per species, levels finest first, one
// patch per finest index and level (patches of a level do not overlap).
__global__ void k_assemble(const MomentSrc *src, const int *level_start, int n_species,
                           int n_levels, int n_f, double *charge, double *J,
                           double *charges_species) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n_f) return;
    double jsum = 0.0;  // EMSolver::J, accumulated over species and levels
    double ch = 0.0;
    for (int s = 0; s < n_species; s++) {
        double cs_ = 0.0;  // charges[s][x]
        for (int l = 0; l < n_levels; l++) {
            double cl = 0.0, jl = 0.0;  // chargeL / currentL of level l
            int a = level_start[s * n_levels + l], b = level_start[s * n_levels + l + 1];
            for (int q = a; q < b; q++) {
                int off = x - src[q].start;
                if (off >= 0 && off < src[q].count) {
                    cl += src[q].chargeR[off];
                    jl += src[q].currentR[off];
                }
            }
            cs_ += cl;
            jsum += jl;
        }
        if (charges_species) charges_species[(long)s * n_f + x] = cs_;
        ch += cs_;
    }
    charge[x] = ch;
    J[x] = jsum;
}

void fill_args(MomArgs &A, const vx_patch *p, const vx_fields *fl) {
    A.p = *p;
    if (fl) A.fl = *fl;
    double m_inv = 1 / p->m;
    A.c1 = m_inv * vx_c_inv;
    A.c2 = 1 / A.c1;
    A.c3 = 1 / 48.0;
    double rtb = (double)p->rtb;
    A.inv_rtb = 1.0 / rtb;
    A.dpq = p->dp * p->q;
    A.jscale = -p->q * p->q / p->m;
    for (int i = 0; i < 3 * MAX_RTB; i++) A.rel[i] = 0.0;
    for (int i = 0; i < p->rtb; i++) {  // Rectangle.cpp:103-110
        double tl = -0.5 + i / rtb;
        double tr = -0.5 + (i + 1.0) / rtb;
        A.rel[3 * i] = (tl + tr) * 0.5;
        A.rel[3 * i + 1] = (tl * tl + tl * tr + tr * tr) / 3.0 - (1.0 / 12);
        A.rel[3 * i + 2] = (tl * tl * tl + tl * tl * tr + tl * tr * tr + tr * tr * tr) * 0.25;
    }
    A.tab = nullptr;
}

}  // namespace

extern "C" {

int vx_energy_patch(vx_ctx *ctx, const vx_patch *p, double *energyR) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (!energyR) return fail(ctx, VX_EINVAL, "energy: null output");
    MomArgs A{};
    fill_args(A, p, nullptr);
    const unsigned nb = (unsigned)((p->n_p * p->rtb + 255) / 256);
    switch (p->rtb) {
        case 1: k_energy<1><<<nb, 256, 0, ctx->stream>>>(A, energyR); break;
        case 2: k_energy<2><<<nb, 256, 0, ctx->stream>>>(A, energyR); break;
        case 4: k_energy<4><<<nb, 256, 0, ctx->stream>>>(A, energyR); break;
        default: return fail(ctx, VX_ENOTSUP, "energy: relativeToBottom %d not in {1,2,4}", p->rtb);
    }
    return check_launch(ctx, "k_energy");
}

}  // extern "C"

static int moments_patch(vx_ctx *ctx, const vx_patch *p, const vx_fields *fl, bool comp) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (p->rtb > MAX_RTB) return fail(ctx, VX_ENOTSUP, "moments: relativeToBottom %d > %d", p->rtb, MAX_RTB);
    if (!fl) return fail(ctx, VX_EINVAL, "moments: null fields");
    if ((rc = validate_span(ctx, p, fl->n_f))) return rc;
    MomArgs A;
    fill_args(A, p, fl);
    const size_t need = 2 * (size_t)tab_stride(p->n_p);

    if (ctx->mom_table_n < need) {
        if (ctx->capturing) return fail(ctx, VX_EINVAL, "moments: table growth while a step graph is recorded");
        if (ctx->mom_table) (void)hipFree(ctx->mom_table);
        ctx->mom_table = nullptr;
        ctx->mom_table_n = 0;
        if (hipMalloc(&ctx->mom_table, sizeof(double) * need) != hipSuccess)
            return fail(ctx, VX_ENOMEM, "moments: logarithm table");
        // no table in place yet (flag 0)
        if (hipMemsetAsync(ctx->mom_table, 0, sizeof(double) * need, ctx->stream) != hipSuccess)
            return fail(ctx, VX_EHIP, "moments: logarithm table clear");
        ctx->mom_table_n = need;
    }
    A.tab = ctx->mom_table;
    switch (p->rtb) {
        case 1: k_moment_table<1><<<2, 1024, 0, ctx->stream>>>(A, ctx->mom_table); break;
        case 2: k_moment_table<2><<<2, 1024, 0, ctx->stream>>>(A, ctx->mom_table); break;
        case 4: k_moment_table<4><<<2, 1024, 0, ctx->stream>>>(A, ctx->mom_table); break;
        default: return fail(ctx, VX_ENOTSUP, "moments: relativeToBottom %d not in {1,2,4}", p->rtb);
    }
    if ((rc = check_launch(ctx, "k_moment_table"))) return rc;
    LaunchTimer lt(ctx, "k_moments");
    const unsigned g1 = (unsigned)((p->n_x + 15) / 16), g2 = (unsigned)((p->n_x + 7) / 8),
                   g4 = (unsigned)((p->n_x + 3) / 4);
    // wavefronts of the launch (4 sub-columns each) against the device's SIMDs;
    // VX_MOMENTS_FORM=split, =pair or =wide pins the form (the ordered forms
    // all give the same bits); the wide form serves relativeToBottom 1 patches
    // without nested cells (the finest level's) and with at least 512 p-cells
    // by default, ordered or compensated
    bool fuse = (size_t)(p->n_x * p->rtb + 3) / 4 <= (size_t)4 * ctx->n_cu;
    // (the wide form runs a column's chunks in lockstep with its workgroup's
    // barriers: for short columns the split and pair forms are faster)
    bool wide = p->rtb == 1 && !p->nested && p->n_p >= 512;
    if (const char *form = std::getenv("VX_MOMENTS_FORM")) {
        if (std::strcmp(form, "split") == 0) fuse = wide = false;
        if (std::strcmp(form, "pair") == 0) {
            fuse = true;
            wide = false;
        }
        if (std::strcmp(form, "wide") == 0) wide = p->rtb == 1 && !p->nested;
    }
    if (wide) {
        const unsigned gw = (unsigned)((p->n_x + WCOL - 1) / WCOL);
        if (comp) k_moments_wide<4, true><<<gw, 64 * WCOL, 0, ctx->stream>>>(A);
        else k_moments_wide<4, false><<<gw, 64 * (WCOL + 1), 0, ctx->stream>>>(A);
        return check_launch(ctx, "k_moments_wide");
    }
    // the pair form: one workgroup of two wavefronts per group of 4 sub-columns
    const unsigned gp1 = (unsigned)((p->n_x + 3) / 4), gp2 = (unsigned)((p->n_x + 1) / 2),
                   gp4 = (unsigned)p->n_x;
    switch (p->rtb * 4 + (comp ? 2 : 0) + (fuse ? 1 : 0)) {
        case 4: k_moments<1, 4, false, false><<<g1, 256, 0, ctx->stream>>>(A); break;
        case 5: k_moments<1, 4, false, false, true><<<gp1, 128, 0, ctx->stream>>>(A); break;
        case 6: case 7: k_moments<1, 4, true, false><<<g1, 256, 0, ctx->stream>>>(A); break;
        case 8: k_moments<2, 2, false, false><<<g2, 256, 0, ctx->stream>>>(A); break;
        case 9: k_moments<2, 2, false, false, true><<<gp2, 128, 0, ctx->stream>>>(A); break;
        case 10: case 11: k_moments<2, 2, true, false><<<g2, 256, 0, ctx->stream>>>(A); break;
        case 16: k_moments_w4<4, 1, false, false><<<g4, 256, 0, ctx->stream>>>(A); break;
        case 17: k_moments<4, 1, false, false, true><<<gp4, 128, 0, ctx->stream>>>(A); break;
        case 18: case 19: k_moments<4, 1, true, false><<<g4, 256, 0, ctx->stream>>>(A); break;
        default: return fail(ctx, VX_ENOTSUP, "moments: relativeToBottom %d not in {1,2,4}", p->rtb);
    }
    return check_launch(ctx, "k_moments");
}

extern "C" {

int vx_moments_patch(vx_ctx *ctx, const vx_patch *p, const vx_fields *fl) {
    return moments_patch(ctx, p, fl, false);
}

int vx_moments_patch_compensated(vx_ctx *ctx, const vx_patch *p, const vx_fields *fl) {
    return moments_patch(ctx, p, fl, true);
}

int vx_moment_plan_create(vx_ctx *ctx, const vx_patch *patches, const vx_moment_layout *lay,
                          int32_t n_f, vx_moment_plan **out) {
    if (!out) return fail(ctx, VX_EINVAL, "null plan pointer");
    *out = nullptr;
    if (!lay || lay->n_species < 1 || lay->n_levels < 1 || n_f < 1)
        return fail(ctx, VX_EINVAL, "bad moment layout");
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "moment plan while a step graph is recorded");
    int ngroups = lay->n_species * lay->n_levels;
    int npatch = lay->level_start[ngroups];
    std::vector<MomentSrc> src(npatch > 0 ? npatch : 1);
    for (int q = 0; q < npatch; q++) {
        src[q].chargeR = patches[q].chargeR;
        src[q].currentR = patches[q].currentR;
        src[q].start = patches[q].x_pos * patches[q].rtb;
        src[q].count = patches[q].n_x * patches[q].rtb;
    }
    size_t sb = sizeof(MomentSrc) * src.size();
    size_t bytes = sb + sizeof(int) * (ngroups + 1);
    vx_moment_plan *pl = new vx_moment_plan;
    pl->n_species = lay->n_species;
    pl->n_levels = lay->n_levels;
    pl->n_f = n_f;
    if (hipMalloc(&pl->tables, bytes) != hipSuccess) {
        delete pl;
        return fail(ctx, VX_ENOMEM, "moment plan tables");
    }
    std::vector<char> host(bytes);
    std::memcpy(host.data(), src.data(), sb);
    std::memcpy(host.data() + sb, lay->level_start, sizeof(int) * (ngroups + 1));
    hipError_t e = hipMemcpy(pl->tables, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(pl->tables);
        delete pl;
        return fail(ctx, VX_EHIP, "moment plan copy: %s", hipGetErrorString(e));
    }
    pl->src_off = 0;
    pl->ls_off = sb;
    *out = pl;
    return VX_OK;
}

void vx_moment_plan_destroy(vx_moment_plan *pl) {
    if (!pl) return;
    (void)hipFree(pl->tables);
    delete pl;
}

int vx_moments_assemble(vx_ctx *ctx, const vx_moment_plan *pl, double *charge, double *J,
                        double *charges_species) {
    if (!pl) return fail(ctx, VX_EINVAL, "null moment plan");
    const MomentSrc *dsrc = (const MomentSrc *)((char *)pl->tables + pl->src_off);
    const int *dls = (const int *)((char *)pl->tables + pl->ls_off);
    k_assemble<<<(unsigned)((pl->n_f + 255) / 256), 256, 0, ctx->stream>>>(
        dsrc, dls, pl->n_species, pl->n_levels, pl->n_f, charge, J, charges_species);
    return check_launch(ctx, "k_assemble");
}

}  // extern "C"
