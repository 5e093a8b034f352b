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
#include <cmath>
#include <cstdlib>
#include <initializer_list>
#include <algorithm>
#include <vector>
#include <tuple>
#include <mutex>
#include <map>

#include "vx_dev.h"
#include "vx_internal.h"

using namespace vxi;
using namespace vxd;

namespace {

// values of one sub-step-0 launch that every patch shares
struct Sub0Coef {
    vx_fields fl;
    int step;
    int ns;          // step + 1 terms in the RK combination
    double a[6];     // RK coefficients times dt (Rectangle.cpp:1396-1517)
    double aSum;     // a[0] + ... + a[ns-1], summed left to right
    double b5[6];    // stage 5's coefficients times dt (stage 4 forms its partial sum)
    double sinv = 1.0;   // mixed precision: 2^-k, the scale of the FP32 copies of f
    float eps32 = 0.0f;  // mixed precision: the WENO epsilon (1e-10) in the scaled units
};

// EMFieldSolver::GetASquared(GetFinestIndex(i)) scaled by q*q (Rectangle.cpp:1280)
__device__ __forceinline__ double face_as(const vx_patch &P, const vx_fields &fl, int i) {
    long fi = (long)P.rtb * (long)(i + P.x_pos);
    long c = fi < 0 ? 0 : (fi > fl.n_f ? fl.n_f : fi);
    return P.q * P.q * fl.a2[c];
}

// Rectangle::GetEfield(i) (Rectangle.cpp:1055-1067); E is periodic on the
// finest grid (EMSolver.cpp:137-154 wraps every stencil index).
__device__ __forceinline__ double patch_E(const vx_patch &p, const vx_fields &fl, int i) {
    return vxd::patch_E(fl.E, fl.n_f, p.rtb, p.x_pos, i);
}

// Sub-step 0 fluxes (Rectangle.cpp:1269-1517 up to the predictor) in
// register-streaming form, for a patch with any edges: a wavefront owns 64
// consecutive p positions, lane l <-> j = jw - 3 + l, and produces the 58
// positions of lanes 3..60 (p neighbours up to 2 lanes away come from DPP lane
// shifts); it marches along x over a strip of TI positions, position ix taking
// rows ix-2 .. ix+1 of f_s, so a strip starts 3 rows early.  Every position
// (i, j) in [-2, n_x+2) x [-2, n_p+2) has one owner, which writes FxDS/FpDS
// there, and FxH_s/FpH_s (FxL0/FpL0 at stage 0) where the reference's loops
// compute them; later stages read the stored FxL0/FpL0 (the coarse-fine flux
// program may have replaced them at interface faces).
// MP: the WENO weights from FP32 copies of f scaled by A.sinv (vxd::weno_mp);
// candidates, fluxes and the RK combination stay FP64.
// Stage 5 needs the earlier stages' fluxes only through its combination: stage
// 4 also writes stage 5's partial sum b_0 F_0 + ... + b_4 F_4 (left to right,
// stage 5's order) into the stage-5 flux planes, which stage 5 reads at each
// position it owns before writing its own fluxes there (positions it does not
// own are only warm-up whose results are discarded); the coarse-fine flux
// program writes the sum at the faces whose fluxes it replaces.
constexpr int S0W = 58;  // positions per wavefront

struct Sub0Tiles {
    int TI, nstrip, nbx;
};

template <int STEP, bool MP>
__device__ __forceinline__ void sub0s_tile(const vx_patch &P, const Sub0Coef &A, const Sub0Tiles &T,
                                           const int tile) {
    const int nx = P.n_x, np = P.n_p, pitch = P.pitch;
    const int lane = threadIdx.x & 63;
    const int bx = tile % T.nbx, by = tile / T.nbx;
    const int strip = by * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (strip >= T.nstrip) return;  // uniform per wavefront
    const int i0 = -2 + strip * T.TI;
    const int i1 = min(i0 + T.TI, nx + 2);  // positions [i0, i1)
    const int jw = -2 + bx * S0W;
    const int j = jw - 3 + lane;
    const bool own_j = lane >= 3 && lane < 3 + S0W && j < np + 2;
    const bool in_j = j >= -2 && j < np + 2;
    const bool jr_ex = j >= -1 && j < np + 1;
    const bool jr_ep = j >= -1 && j < np + 2;
    const bool jr_fp = j >= 0 && j < np + 1;
    const double dpcc = P.dp_inv * P.cc;
    const double ccdx = P.cc * P.dx_inv;
    const double w3 = 1 / 48.0;
    const double pj = momentum(P.pmin, P.dp, j, P.p_pos);
    const double *f1 = P.f[1];
    auto ld = [&](int i) { return (i >= -2 && i < nx + 2 && in_j) ? f1[pidx(pitch, i, j)] : 0.0; };

    double fs_m3 = 0, fs_m2 = 0, fs_m1 = 0, fs_0 = 0;
    float gs_m3 = 0, gs_m2 = 0, gs_m1 = 0, gs_0 = 0;  // MP: scaled FP32 copies
    double ep_m2 = 0, ep_m1 = 0, ep_0 = 0, fp_m2 = 0, fp_m1 = 0, fp_0 = 0;
    const int r0 = i0 - 2;
    double as_m1 = face_as(P, A.fl, r0 - 1), as_0 = face_as(P, A.fl, r0);
    double G_m1 = gamma_(pj, as_m1, P.mc2inv);
    double G_0 = (as_0 == as_m1) ? G_m1 : gamma_(pj, as_0, P.mc2inv);
    double pfs_a = ld(r0), pfs_b = ld(r0 + 1);
    // stored planes read: those of the earlier stages, at stage 5 only its
    // partial sum (in plane 5)
    constexpr int NH = STEP == 5 ? 1 : (STEP > 0 ? STEP : 1);
    double hx[NH], hp[NH], lx = 0.0, lp = 0.0;
    auto load_h = [&](int pos) {  // stored planes at position (pos, j)
        const bool v = pos >= -2 && pos < nx + 2 && in_j;
#pragma unroll
        for (int k = 0; k < NH; k++) {
            const int plane = STEP == 5 ? 5 : k;
            hx[k] = (k < STEP && v) ? P.FxH[plane][pidx(pitch, pos, j)] : 0.0;
            hp[k] = (k < STEP && v) ? P.FpH[plane][pidx(pitch, pos, j)] : 0.0;
        }
        if (STEP > 0) {
            lx = v ? P.FxL0[pidx(pitch, pos, j)] : 0.0;
            lp = v ? P.FpL0[pidx(pitch, pos, j)] : 0.0;
        }
    };
    load_h(r0 - 1);

    for (int r = r0; r <= i1; r++) {
        fs_0 = pfs_a;
        if (MP) gs_0 = (float)(fs_0 * A.sinv);
        pfs_a = pfs_b;
        pfs_b = ld(r + 2);
        double chx[NH], chp[NH];
#pragma unroll
        for (int k = 0; k < NH; k++) {
            chx[k] = hx[k];
            chp[k] = hp[k];
        }
        const double clx = lx, clp = lp;
        load_h(r);
        const double as_p1 = face_as(P, A.fl, r + 1);
        const double G_p1 = (as_p1 == as_0) ? G_0 : gamma_(pj, as_p1, P.mc2inv);

        // x-face ix = r - 1 (Rectangle.cpp:1278-1293, 1313-1352)
        const int ix = r - 1;
        const double Gd = shdn1(G_m1);
        const double ex = (ix >= -1 && ix < nx + 2 && jr_ex) ? dpcc * (Gd - G_m1) : 0.0;
        const bool vFx = ix >= 0 && ix < nx + 1 && jr_ex;
        double fx = 0.0;
        if (vFx) {
            if (MP) fx = weno_mp(fs_m3, fs_m2, fs_m1, fs_0, gs_m3, gs_m2, gs_m1, gs_0, A.eps32, ex > 0.0);
            else fx = weno(fs_m3, fs_m2, fs_m1, fs_0, ex > 0.0);
        }
        const double ex_u = shdn1(ex), ex_d = shup1(ex), fx_u = shdn1(fx), fx_d = shup1(fx);
        const double fxh = vFx ? P.dx_inv * (fx * ex + w3 * (fx_u - fx_d) * (ex_u - ex_d)) : 0.0;
        const double fxl = vFx ? P.dx_inv * ((ex > 0.0 ? fs_m2 : fs_m1) * ex) : 0.0;

        // column r (Rectangle.cpp:1294-1312)
        const bool cr = r >= -1 && r < nx + 1;
        const double qE = cr ? P.q * patch_E(P, A.fl, r) : 0.0;
        ep_0 = (cr && jr_ep) ? qE - ccdx * (G_p1 - G_0) : 0.0;
        const double s1 = shup1(fs_0), s2 = shup1(s1), s3 = shdn1(fs_0);
        fp_0 = 0.0;
        if (MP) {
            const float t1 = shup1f(gs_0), t2 = shup1f(t1), t3 = shdn1f(gs_0);
            if (cr && jr_fp) fp_0 = weno_mp(s2, s1, fs_0, s3, t2, t1, gs_0, t3, A.eps32, ep_0 > 0.0);
        } else if (cr && jr_fp) {
            fp_0 = weno(s2, s1, fs_0, s3, ep_0 > 0.0);
        }

        // column ix = r - 1 (Rectangle.cpp:1353-1394)
        const bool vFp = ix >= -1 && ix < nx + 1 && jr_fp;
        const double fph = vFp ? P.dp_inv * (fp_m1 * ep_m1 + w3 * (fp_0 - fp_m2) * (ep_0 - ep_m2)) : 0.0;
        const double fsd = shup1(fs_m1);  // f(ix, j-1)
        const double fpl = vFp ? P.dp_inv * ((ep_m1 > 0.0 ? fsd : fs_m1) * ep_m1) : 0.0;

        // position (ix, j): stores and the RK combination (Rectangle.cpp:1396-1517)
        if (ix >= i0 && ix < i1 && own_j) {
            const long id = pidx(pitch, ix, j);
            // FxH_s/FpH_s are read again only by the later stages: streamed
            // past the caches, which keep what the next kernels read
            if (vFx) {
                __builtin_nontemporal_store(fxh, &P.FxH[STEP][id]);
                if (STEP == 0) P.FxL0[id] = fxl;
            }
            if (vFp) {
                __builtin_nontemporal_store(fph, &P.FpH[STEP][id]);
                if (STEP == 0) P.FpL0[id] = fpl;
            }
            if (STEP == 4) {  // stage 5's partial sum, in its order
                double qx = A.b5[0] * chx[0], qp = A.b5[0] * chp[0];
#pragma unroll
                for (int k = 1; k < NH; k++) {
                    qx = qx + A.b5[k] * chx[k];
                    qp = qp + A.b5[k] * chp[k];
                }
                if (vFx) __builtin_nontemporal_store(qx + A.b5[4] * fxh, &P.FxH[5][id]);
                if (vFp) __builtin_nontemporal_store(qp + A.b5[4] * fph, &P.FpH[5][id]);
            }
            const double xLS = A.aSum * (STEP == 0 ? fxl : clx);
            const double pLS = A.aSum * (STEP == 0 ? fpl : clp);
            double sx, sp;
            if (STEP == 5) {  // the partial sum stage 4 formed (+0 where not stored)
                sx = chx[0] + A.a[5] * fxh;
                sp = chp[0] + A.a[5] * fph;
            } else {
                sx = A.a[0] * (STEP == 0 ? fxh : chx[0]);
                sp = A.a[0] * (STEP == 0 ? fph : chp[0]);
#pragma unroll
                for (int k = 1; k <= STEP; k++) {
                    sx = sx + A.a[k] * (k == STEP ? fxh : chx[k < NH ? k : 0]);
                    sp = sp + A.a[k] * (k == STEP ? fph : chp[k < NH ? k : 0]);
                }
            }
            P.FxDS[id] = sx - xLS;
            P.FpDS[id] = sp - pLS;
        }

        fs_m3 = fs_m2; fs_m2 = fs_m1; fs_m1 = fs_0;
        gs_m3 = gs_m2; gs_m2 = gs_m1; gs_m1 = gs_0;
        as_0 = as_p1; G_m1 = G_0; G_0 = G_p1;
        ep_m2 = ep_m1; ep_m1 = ep_0; fp_m2 = fp_m1; fp_m1 = fp_0;
    }
}

template <int STEP>
__global__ __launch_bounds__(256) void k_sub0s(vx_patch P, Sub0Coef A, Sub0Tiles T) {
    sub0s_tile<STEP, false>(P, A, T, xcd_order((int)blockIdx.x, (int)gridDim.x));
}

// The first patch k of the work list whose workgroups start at or before w
// (uniform per workgroup).
template <int KIND>
__device__ __forceinline__ int batch_patch(const vx_batch_info *bi, int n, int w) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bi[mid].start[KIND] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int STEP, bool MP>
__global__ __launch_bounds__(256) void k_sub0s_batch(const vx_patch *tab, const vx_batch_info *bi, int n,
                                                     Sub0Coef A) {
    const int w = xcd_chunk_order<8>((int)blockIdx.x, (int)gridDim.x);
    const int k = batch_patch<0>(bi, n, w);
    const Sub0Tiles T{bi[k].TI, bi[k].nstrip, bi[k].nbx};
    sub0s_tile<STEP, MP>(tab[k], A, T, w - bi[k].start[0]);
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

__device__ __forceinline__ void lo_predictor(const vx_patch &P, double aSum, int bx, int by) {
    int j = bx * 64 + (threadIdx.x & 63);
    int i = by * 4 + (threadIdx.x >> 6);
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

// Zalesak limiter, both passes in one launch (Rectangle.cpp:1538-1594): R+- on
// [-1, n_x+1) x [-1, n_p+1), then C on the NS range [-2, n_x+2) x [-2, n_p+2):
// 1 everywhere, the limiter formula on interior faces i in [1, n_x), j in
// [0, n_p).  Register-streaming form: a wavefront owns 64 consecutive p
// positions, lane l <-> j = jw - 2 + l, and produces lanes 2..57 (56
// positions, so that each wavefront's own row segments start on a 64-byte
// boundary of the pitched row; p neighbours come from DPP lane shifts, and the
// corrector's C at (i, j+1) of lane 57 from lane 58, whose R needs lane 59).  It marches along x over
// a strip of C rows [i0, i1): at row i it forms R of row i from rows i-1..i+1
// of f^n, f^td and the antidiffusive fluxes (held in registers, one new row
// loaded per step), and C of row i from R of rows i and i-1.  R is stored
// where it is formed inside the strip (the cross-patch C sync reads it); the
// first row of a strip (i0 - 1) only warms the pipeline.  MP: the ratios Q / P
// in FP32 on the values scaled by sinv (vxd::zalesak_mp).
// COR: the corrector too (Rectangle.cpp:1596-1612, vxi::corrected_cell with
// C, f^td and the fluxes from registers): row i - 1 is corrected at step i,
// once C of row i is formed, so a strip takes one more step (row i1, whose R
// and C are formed but not stored).  The values use C before the cross-patch
// C synchronisation; cells next to a face that the synchronisation changes
// are corrected again after it (VX_OP_CORRECT).
constexpr int LIMW = 56;  // positions per limiter wavefront (7 x 64 bytes)

struct LimTiles {
    int TI, nstrip, nbx;
};

template <bool MP, bool COR>
__device__ __forceinline__ void limiter_stream(const vx_patch &P, const LimTiles &T, int tile, double sinv) {
    const int nx = P.n_x, np = P.n_p, pitch = P.pitch;
    const int lane = threadIdx.x & 63;
    const int bx = tile % T.nbx, by = tile / T.nbx;
    const int strip = by * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (strip >= T.nstrip) return;  // uniform per wavefront
    const int i0 = -2 + strip * T.TI;
    const int i1 = min(i0 + T.TI, nx + 2);  // C rows [i0, i1)
    const int j = -2 + bx * LIMW - 2 + lane;
    const bool own = lane >= 2 && lane < 2 + LIMW && j < np + 2;
    const bool in_j = j >= -2 && j < np + 2;
    const bool r_j = j >= -1 && j < np + 1;   // R's p range
    const bool c_j = j >= 0 && j < np;        // interior faces in p
    const double *f0 = P.f[0], *f2 = P.f[2], *FxDS = P.FxDS, *FpDS = P.FpDS;
    auto ld = [&](const double *pl, int i) {
        return (i >= -2 && i < nx + 2 && in_j) ? pl[pidx(pitch, i, j)] : 0.0;
    };
    // at row i: rows i-1, i, i+1 of f^n and f^td, rows i, i+1 of FxDS and
    // FpDS in registers; rows i+2 (q) and i+3 (issued at the top of the step)
    // in flight
    const int r0 = i0 - 1;
    double a0_m = ld(f0, r0 - 1), a0_c = ld(f0, r0), a2_m = ld(f2, r0 - 1), a2_c = ld(f2, r0);
    double x_c = ld(FxDS, r0), p_c = ld(FpDS, r0);
    double a0_p = ld(f0, r0 + 1), a2_p = ld(f2, r0 + 1), x_p = ld(FxDS, r0 + 1), p_p = ld(FpDS, r0 + 1);
    double a0_q = ld(f0, r0 + 2), a2_q = ld(f2, r0 + 2), x_q = ld(FxDS, r0 + 2), p_q = ld(FpDS, r0 + 2);
    double Rp_m = 0.0, Rm_m = 0.0;  // R of row i - 1
    // COR: C, FxDS and FpDS of row i - 1
    double cx_m = 1.0, cp_m = 1.0, x_m = 0.0, p_m = 0.0;
    double *const f1 = P.f[1];
    const int iend = COR ? i1 + 1 : i1;
    for (int i = r0; i < iend; i++) {
        const double na0 = ld(f0, i + 3), na2 = ld(f2, i + 3), nx_ = ld(FxDS, i + 3), np_ = ld(FpDS, i + 3);
        // R+- of (i, j) (Rectangle.cpp:1538-1579)
        const double p_u = shdn1(p_c);  // FpDS(i, j+1)
        double rp = 0.0, rm = 0.0;
        {
            const double Pp = valmax(0.0, x_c) - valmin(0.0, x_p) + valmax(0.0, p_c) - valmin(0.0, p_u);
            const double Pm = valmax(0.0, x_p) - valmin(0.0, x_c) + valmax(0.0, p_u) - valmin(0.0, p_c);
            const double d0 = shdn1(a0_c), d2 = shdn1(a2_c);  // (i, j+1)
            const double e0 = shup1(a0_c), e2 = shup1(a2_c);  // (i, j-1)
            const double wMax = valmax(valmax(a0_c, a2_c), valmax(valmax(a0_p, a2_p), valmax(valmax(a0_m, a2_m),
                                valmax(valmax(d0, d2), valmax(e0, e2)))));
            const double wMin = valmin(valmin(a0_c, a2_c), valmin(valmin(a0_p, a2_p), valmin(valmin(a0_m, a2_m),
                                valmin(valmin(d0, d2), valmin(e0, e2)))));
            const double Qm = -wMin + a2_c;
            const double Qp = wMax - a2_c;
            if (i >= -1 && i < nx + 1 && r_j) {
                if (MP) {
                    rp = Pp > 0.0 ? zalesak_mp(Qp, Pp, sinv) : 0.0;
                    rm = Pm > 0.0 ? zalesak_mp(Qm, Pm, sinv) : 0.0;
                } else {
                    rp = Pp > 0.0 ? valmin(1.0, Qp / Pp) : 0.0;
                    rm = Pm > 0.0 ? valmin(1.0, Qm / Pm) : 0.0;
                }
            }
        }
        double cx = 1.0, cp = 1.0;
        if (i >= i0) {
            const long idx = pidx(pitch, i, j);
            if (own && i < i1 && i >= -1 && i < nx + 1 && r_j) {
                // only the cross-patch C sync reads R again (at patch edges):
                // streamed past the caches, which keep C and the fluxes that
                // the corrector reads next
                __builtin_nontemporal_store(rp, &P.Rp[idx]);
                __builtin_nontemporal_store(rm, &P.Rm[idx]);
            }
            // C of (i, j): R(i, j), R(i-1, j), R(i, j-1) (Rectangle.cpp:1581-1594)
            const double Rp_l = shup1(rp), Rm_l = shup1(rm);
            if (i >= 1 && i < nx && c_j) {
                cx = x_c > 0.0 ? valmin(rp, Rm_m) : valmin(Rp_m, rm);
                cp = p_c > 0.0 ? valmin(rp, Rm_l) : valmin(Rp_l, rm);
            }
            if (own && i < i1) {
                P.Cx[idx] = cx;
                P.Cp[idx] = cp;
            }
        }
        if (COR) {
            // cell (c, j) = (i - 1, j): the terms of corrected_cell in its order
            const int c = i - 1;
            const double cp_u = shdn1(cp_m), fp_u = shdn1(p_m);  // Cp, FpDS at (c, j + 1)
            if (own && c >= i0 && c >= 0 && c < nx && j >= 0 && j < np) {
                double t = a2_m;
                if (in_update_loop(P, c, j)) {
                    t += cx_m * x_m;
                    t += cp_m * p_m;
                }
                if (in_update_loop(P, c, j + 1)) t -= cp_u * fp_u;
                if (in_update_loop(P, c + 1, j)) t -= cx * x_c;
                f1[pidx(pitch, c, j)] = t;
            }
            cx_m = cx;
            cp_m = cp;
            x_m = x_c;
            p_m = p_c;
        }
        Rp_m = rp;
        Rm_m = rm;
        a0_m = a0_c; a0_c = a0_p; a0_p = a0_q; a0_q = na0;
        a2_m = a2_c; a2_c = a2_p; a2_p = a2_q; a2_q = na2;
        x_c = x_p; x_p = x_q; x_q = nx_;
        p_c = p_p; p_p = p_q; p_q = np_;
    }
}

// f1 = f2 + limited antidiffusive fluxes on interior cells
// (Rectangle.cpp:1596-1612, vxi::corrected_cell), one thread per cell.
__device__ __forceinline__ void corrector(const vx_patch &P, int bx, int by) {
    int j = bx * 64 + (threadIdx.x & 63);
    int i = by * 4 + (threadIdx.x >> 6);
    if (i >= P.n_x || j >= P.n_p) return;
    P.f[1][pidx(P.pitch, i, j)] = corrected_cell(P, i, j);
}

__global__ __launch_bounds__(256) void k_lo_predictor(vx_patch P, double aSum) {
    lo_predictor(P, aSum, (int)blockIdx.x, (int)blockIdx.y);
}
__global__ __launch_bounds__(256) void k_limiter(vx_patch P, LimTiles T) {
    limiter_stream<false, false>(P, T, xcd_order((int)blockIdx.x, (int)gridDim.x), 1.0);
}
__global__ __launch_bounds__(256) void k_corrector(vx_patch P) { corrector(P, (int)blockIdx.x, (int)blockIdx.y); }

// Work-list forms: workgroup w of the launch is workgroup w - start of patch k,
// laid out as that patch's own 2-D grid of 64 x 4 cells (grid_cells).
template <int KIND, int EXTRA>
__device__ __forceinline__ void batch_cell(const vx_batch_info *bi, int n, int &k, int &bx, int &by,
                                           const vx_patch *tab) {
    const int w = xcd_order((int)blockIdx.x, (int)gridDim.x);
    k = batch_patch<KIND>(bi, n, w);
    const int local = w - bi[k].start[KIND];
    const int gx = (tab[k].n_p + EXTRA + 63) / 64;
    bx = local % gx;
    by = local / gx;
}

__global__ __launch_bounds__(256) void k_lo_predictor_batch(const vx_patch *tab, const vx_batch_info *bi, int n,
                                                            double aSum) {
    int k, bx, by;
    batch_cell<1, 0>(bi, n, k, bx, by, tab);
    lo_predictor(tab[k], aSum, bx, by);
}
template <bool MP, bool COR>
__global__ __launch_bounds__(256) void k_limiter_batch(const vx_patch *tab, const vx_batch_info *bi, int n,
                                                       double sinv) {
    const int w = xcd_chunk_order<8>((int)blockIdx.x, (int)gridDim.x);
    const int k = batch_patch<2>(bi, n, w);
    const LimTiles T{bi[k].lTI, bi[k].lnstrip, bi[k].lnbx};
    limiter_stream<MP, COR>(tab[k], T, w - bi[k].start[2], sinv);
}
__global__ __launch_bounds__(256) void k_corrector_batch(const vx_patch *tab, const vx_batch_info *bi, int n) {
    int k, bx, by;
    batch_cell<1, 0>(bi, n, k, bx, by, tab);
    corrector(tab[k], bx, by);
}
// f0 <- f1 over the whole plane of every patch (Rectangle.cpp:1614-1622)
__global__ __launch_bounds__(256) void k_copy_plane_batch(const vx_patch *tab, const vx_batch_info *bi, int n) {
    const int k = batch_patch<4>(bi, n, (int)blockIdx.x);
    const long e = (long)((int)blockIdx.x - bi[k].start[4]) * 256 + threadIdx.x;
    const vx_patch &P = tab[k];
    if (e < (long)(P.n_x + 4) * P.pitch) P.f[0][e] = P.f[1][e];
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

// Strip height of a streaming kernel over n positions in x with nbx columns
// of wavefronts: pol > 0 is the height itself (at most the positions), pol =
// -W the tallest of `tall` that still gives W wavefronts for the patch.
int strip_rows(int pol, int n, int nbx, std::initializer_list<int> tall, int floor_ti) {
    if (pol > 0) return std::min(pol, n);
    int ti = floor_ti;
    for (int t : tall)
        if ((long)nbx * ((n + t - 1) / t) >= -pol) {
            ti = t;
            break;
        }
    return ti;
}

// strips of the streaming sub-step-0 kernel (default: the tallest that still
// gives 2048 wavefronts)
Sub0Tiles sub0_tiles(const vx_patch &p, int pol = -2048) {
    Sub0Tiles T;
    T.nbx = (p.n_p + 4 + S0W - 1) / S0W;
    T.TI = strip_rows(pol, p.n_x + 4, T.nbx, {64, 32, 16, 8}, 4);
    T.nstrip = (p.n_x + 4 + T.TI - 1) / T.TI;
    return T;
}

// strips of the streaming limiter (default: the tallest that still gives
// 4096 wavefronts)
LimTiles lim_tiles(const vx_patch &p, int pol = -4096) {
    LimTiles T;
    T.nbx = (p.n_p + 4 + LIMW - 1) / LIMW;
    T.TI = strip_rows(pol, p.n_x + 4, T.nbx, {32, 16, 8}, 4);
    T.nstrip = (p.n_x + 4 + T.TI - 1) / T.TI;
    return T;
}

int table_rebuild(vx_ctx *ctx, vx_patch_table *t);

// the work list of a patch table (built once: a table's shapes never change)
int table_batch(vx_ctx *ctx, vx_patch_table *t, int n_f) {
    if (!t || t->n < 1) return fail(ctx, VX_EINVAL, "batched sub-step: empty patch table");
    for (const vx_patch &p : t->host) {
        int rc = validate_span(ctx, &p, n_f);
        if (rc) return rc;
    }
    if (t->batch_d) return VX_OK;
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "batched sub-step: work list built while recording");
    return table_rebuild(ctx, t);
}

// (Re)build the work list of a table for its current strip policies; the
// device copy is rewritten in place once the stream has drained.
int table_rebuild(vx_ctx *ctx, vx_patch_table *t) {
    t->batch.assign(t->n + 1, vx_batch_info{});
    int run[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k <= t->n; k++) {
        vx_batch_info &b = t->batch[k];
        for (int q = 0; q < 5; q++) b.start[q] = run[q];
        if (k == t->n) break;
        const vx_patch &p = t->host[k];
        const Sub0Tiles T = sub0_tiles(p, t->sub0_pol);
        b.TI = T.TI;
        b.nstrip = T.nstrip;
        b.nbx = T.nbx;
        const LimTiles TL = lim_tiles(p, t->lim_pol);
        b.lTI = TL.TI;
        b.lnstrip = TL.nstrip;
        b.lnbx = TL.nbx;
        dim3 g0 = grid_cells(p.n_x, p.n_p, 0, 0), g4 = grid_cells(p.n_x, p.n_p, 4, 4);
        run[0] += T.nbx * ((T.nstrip + 3) / 4);
        run[1] += (int)(g0.x * g0.y);
        run[2] += TL.nbx * ((TL.nstrip + 3) / 4);
        run[3] += (int)(g4.x * g4.y);
        run[4] += (int)(((long)(p.n_x + 4) * p.pitch + 255) / 256);
    }
    const size_t bytes = sizeof(vx_batch_info) * t->batch.size();
    int rc = VX_OK;
    if (!t->batch_d) rc = check_hip(ctx, hipMalloc((void **)&t->batch_d, bytes), "work list");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "work list");
    if (!rc) rc = check_hip(ctx, hipMemcpy(t->batch_d, t->batch.data(), bytes, hipMemcpyHostToDevice), "work list");
    return rc;
}

// The strip policies of a table's batched flux (kind 0) and limiter (kind 1)
// launches, chosen by timing on the table's first launch of each kind: every
// candidate policy runs the launch three times (the launches are idempotent:
// they read f, the stored fluxes of earlier stages and f^n, and write only
// their own outputs, the same values under any policy), the fastest is kept
// (a later candidate must be 3 % faster), and the choice is kept per device
// for tables of the same patch shapes.
struct PolKey {
    int device, kind, mp;
    std::vector<int> shapes;
    bool operator<(const PolKey &o) const {
        return std::tie(device, kind, mp, shapes) < std::tie(o.device, o.kind, o.mp, o.shapes);
    }
};
std::mutex g_pol_mu;
std::map<PolKey, int> g_pol;

template <class Launch>
int tune_policy(vx_ctx *ctx, vx_patch_table *t, int kind, bool mp, const std::vector<int> &cands, Launch launch) {
    int &pol = kind == 0 ? t->sub0_pol : t->lim_pol;
    // VX_AMR_POLICY="sub0,limiter" pins both policies (a strip height, or
    // minus the wavefront count to reach), e.g. to profile a run without the
    // timing launches or to check that the policy changes no result
    if (const char *env = std::getenv("VX_AMR_POLICY")) {
        int v[2] = {0, 0}, n = 0;
        for (const char *q = env; *q && n < 2;) {
            v[n++] = std::atoi(q);
            while (*q && *q != ',') q++;
            if (*q == ',') q++;
        }
        if (n == 2 && v[kind] != 0 && v[kind] <= 1024 && v[kind] >= -(1 << 24)) {
            if (pol == v[kind]) return VX_OK;
            pol = v[kind];
            return table_rebuild(ctx, t);
        }
    }
    PolKey key{ctx->device, kind, (int)mp, {}};
    for (const vx_patch &p : t->host) {
        key.shapes.push_back(p.n_x);
        key.shapes.push_back(p.n_p);
    }
    {
        std::lock_guard<std::mutex> g(g_pol_mu);
        auto it = g_pol.find(key);
        if (it != g_pol.end()) {
            if (pol == it->second) return VX_OK;
            pol = it->second;
            return table_rebuild(ctx, t);
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = pol, rc = VX_OK;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        bool failed = false;
        // one launch of policy c, timed on its own
        auto timed = [&](int c) {
            float tm = 1e30f;
            pol = c;
            if ((rc = table_rebuild(ctx, t))) return tm;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) failed = true;
            if (!failed) launch();
            if (!failed && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&tm, e0, e1) != hipSuccess))
                failed = true;
            return tm;
        };
        // one warm-up launch, one timed launch of each candidate, then the two
        // fastest timed twice more, keeping their fastest launch; among those,
        // in candidate order, a later one must be more than 3 % faster to win
        pol = cands.empty() ? pol : cands[0];
        if (!(rc = table_rebuild(ctx, t))) launch();  // warm
        std::vector<float> ms(cands.size(), 1e30f);
        for (size_t k = 0; k < cands.size() && !failed && !rc; k++) ms[k] = timed(cands[k]);
        if (!failed && !rc && !cands.empty()) {
            std::vector<int> order(cands.size());
            for (size_t k = 0; k < cands.size(); k++) order[k] = (int)k;
            std::sort(order.begin(), order.end(), [&](int a, int b) { return ms[a] < ms[b]; });
            std::vector<int> fin(order.begin(), order.begin() + std::min<size_t>(2, order.size()));
            for (int k : fin)
                for (int rep = 0; rep < 2 && !failed && !rc; rep++) ms[k] = std::min(ms[k], timed(cands[k]));
            if (!failed && !rc) {
                std::sort(fin.begin(), fin.end());  // candidate order
                float best_ms = 1e30f;
                for (int k : fin)
                    if (ms[k] < 0.97f * best_ms) {
                        best_ms = ms[k];
                        best = cands[k];
                    }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    if ((rc = check_launch(ctx, kind == 0 ? "k_sub0s_batch (tuning)" : "k_limiter_batch (tuning)"))) return rc;
    pol = best;
    {
        std::lock_guard<std::mutex> g(g_pol_mu);
        g_pol[key] = best;
    }
    return table_rebuild(ctx, t);
}

int mp_scale(vx_ctx *ctx, double f_scale, double &sinv, float &eps32, const char *what) {
    if (!(f_scale > 0.0) || !std::isfinite(f_scale))
        return fail(ctx, VX_EINVAL, "%s: f scale must be positive and finite", what);
    sinv = std::ldexp(1.0, -std::ilogb(f_scale));
    eps32 = (float)(1.0e-10 * sinv * sinv);
    return VX_OK;
}

template <bool MP>
void launch_sub0s_batch(vx_ctx *ctx, vx_patch_table *t, int step, const Sub0Coef &A) {
    dim3 grid((unsigned)t->batch[t->n].start[0]);
    LaunchTimer lt(ctx, "k_sub0");
    switch (step) {
        case 0: k_sub0s_batch<0, MP><<<grid, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, A); break;
        case 1: k_sub0s_batch<1, MP><<<grid, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, A); break;
        case 2: k_sub0s_batch<2, MP><<<grid, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, A); break;
        case 3: k_sub0s_batch<3, MP><<<grid, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, A); break;
        case 4: k_sub0s_batch<4, MP><<<grid, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, A); break;
        default: k_sub0s_batch<5, MP><<<grid, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, A); break;
    }
}

int sub0_fluxes_table(vx_ctx *ctx, vx_patch_table *t, int step, double dt, const vx_fields *fl, bool mp,
                             double f_scale) {
    if (step < 0 || step > 5 || !fl) return fail(ctx, VX_EINVAL, "sub0 table: bad step %d", step);
    Sub0Coef A;
    if (mp) {
        int rc = mp_scale(ctx, f_scale, A.sinv, A.eps32, "sub0 table (mixed)");
        if (rc) return rc;
    }
    int rc = table_batch(ctx, t, fl->n_f);
    if (rc) return rc;
    A.fl = *fl;
    A.step = step;
    rk_coefficients(step, dt, A.a, A.ns, A.aSum);
    for (int k = 0; k < 6; k++) A.b5[k] = vx::rk_a[5][k] * dt;  // as stage 5 forms its a[k]
    if (!t->sub0_tuned[mp] && !ctx->capturing) {
        rc = tune_policy(ctx, t, 0, mp, {-2048, 128, 64, 32, 16, 8}, [&] {
            if (mp) launch_sub0s_batch<true>(ctx, t, step, A);
            else launch_sub0s_batch<false>(ctx, t, step, A);
        });
        if (rc) return rc;
        t->sub0_tuned[mp] = true;
    }
    if (mp) launch_sub0s_batch<true>(ctx, t, step, A);
    else launch_sub0s_batch<false>(ctx, t, step, A);
    return check_launch(ctx, "k_sub0s_batch");
}

}  // namespace

extern "C" {

int vx_fct_sub0_fluxes(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (step < 0 || step > 5 || !fl) return fail(ctx, VX_EINVAL, "sub0: bad step %d", step);
    if ((rc = validate_span(ctx, p, fl->n_f))) return rc;
    Sub0Coef A;
    A.fl = *fl;
    A.step = step;
    rk_coefficients(step, dt, A.a, A.ns, A.aSum);
    for (int k = 0; k < 6; k++) A.b5[k] = vx::rk_a[5][k] * dt;  // as stage 5 forms its a[k]
    const Sub0Tiles T = sub0_tiles(*p);
    dim3 grid((unsigned)(T.nbx * ((T.nstrip + 3) / 4)));
    {
        LaunchTimer lt(ctx, "k_sub0");
        switch (step) {
            case 0: k_sub0s<0><<<grid, 256, 0, ctx->stream>>>(*p, A, T); break;
            case 1: k_sub0s<1><<<grid, 256, 0, ctx->stream>>>(*p, A, T); break;
            case 2: k_sub0s<2><<<grid, 256, 0, ctx->stream>>>(*p, A, T); break;
            case 3: k_sub0s<3><<<grid, 256, 0, ctx->stream>>>(*p, A, T); break;
            case 4: k_sub0s<4><<<grid, 256, 0, ctx->stream>>>(*p, A, T); break;
            default: k_sub0s<5><<<grid, 256, 0, ctx->stream>>>(*p, A, T); break;
        }
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
    const LimTiles T = lim_tiles(*p);
    k_limiter<<<(unsigned)(T.nbx * ((T.nstrip + 3) / 4)), 256, 0, ctx->stream>>>(*p, T);
    return check_launch(ctx, "k_limiter");
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


// ---- work-list (batched) forms: one launch per sub-step for every patch of a table
int vx_fct_sub0_fluxes_table(vx_ctx *ctx, vx_patch_table *t, int step, double dt, const vx_fields *fl) {
    return sub0_fluxes_table(ctx, t, step, dt, fl, false, 0.0);
}

int vx_fct_sub0_fluxes_table_mp(vx_ctx *ctx, vx_patch_table *t, int step, double dt, const vx_fields *fl,
                                double f_scale) {
    return sub0_fluxes_table(ctx, t, step, dt, fl, true, f_scale);
}

int vx_fct_sub0_predictor_table(vx_ctx *ctx, vx_patch_table *t, int step, double dt, int32_t n_f) {
    if (step < 0 || step > 5) return fail(ctx, VX_EINVAL, "predictor table: bad step %d", step);
    int rc = table_batch(ctx, t, n_f);
    if (rc) return rc;
    double a[6], aSum;
    int ns;
    rk_coefficients(step, dt, a, ns, aSum);
    k_lo_predictor_batch<<<(unsigned)t->batch[t->n].start[1], 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, aSum);
    return check_launch(ctx, "k_lo_predictor_batch");
}

// cor: the limiter launch also corrects every interior cell (k_limiter_batch<MP, true>)
static int sub1_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f, bool mp, double sinv, bool cor) {
    int rc = table_batch(ctx, t, n_f);
    if (rc) return rc;
    auto launch = [&] {
        // the workgroup count follows the strip policy, which tuning changes
        const unsigned nwg = (unsigned)t->batch[t->n].start[2];
        if (mp && cor) k_limiter_batch<true, true><<<nwg, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, sinv);
        else if (mp) k_limiter_batch<true, false><<<nwg, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, sinv);
        else if (cor) k_limiter_batch<false, true><<<nwg, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, sinv);
        else k_limiter_batch<false, false><<<nwg, 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n, sinv);
    };
    if (!t->lim_tuned[mp] && !ctx->capturing) {
        if ((rc = tune_policy(ctx, t, 1, mp, {-4096, 64, 32, 16, 8, 4}, launch))) return rc;
        t->lim_tuned[mp] = true;
    }
    launch();
    return check_launch(ctx, "k_limiter_batch");
}

int vx_fct_sub1_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f) {
    return sub1_table(ctx, t, n_f, false, 1.0, false);
}

int vx_fct_sub1_table_mp(vx_ctx *ctx, vx_patch_table *t, int32_t n_f, double f_scale) {
    double sinv;
    float eps32;
    int rc = mp_scale(ctx, f_scale, sinv, eps32, "sub1 table (mixed)");
    if (rc) return rc;
    return sub1_table(ctx, t, n_f, true, sinv, false);
}

int vx_fct_sub12_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f) {
    return sub1_table(ctx, t, n_f, false, 1.0, true);
}

int vx_fct_sub12_table_mp(vx_ctx *ctx, vx_patch_table *t, int32_t n_f, double f_scale) {
    double sinv;
    float eps32;
    int rc = mp_scale(ctx, f_scale, sinv, eps32, "sub12 table (mixed)");
    if (rc) return rc;
    return sub1_table(ctx, t, n_f, true, sinv, true);
}

int vx_fct_sub2_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f) {
    int rc = table_batch(ctx, t, n_f);
    if (rc) return rc;
    k_corrector_batch<<<(unsigned)t->batch[t->n].start[1], 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n);
    return check_launch(ctx, "k_corrector_batch");
}

int vx_fct_sub3_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f) {
    int rc = table_batch(ctx, t, n_f);
    if (rc) return rc;
    k_copy_plane_batch<<<(unsigned)t->batch[t->n].start[4], 256, 0, ctx->stream>>>(t->d, t->batch_d, t->n);
    return check_launch(ctx, "k_copy_plane_batch");
}

}  // extern "C"
