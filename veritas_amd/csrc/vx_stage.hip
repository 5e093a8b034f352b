// Fused FCT stage for one patch whose four edges are domain walls (a single
// unrefined level): Rectangle::FCTTimeStep sub-steps 0-2 with the zero ghost
// fills between them (Mesh::Advance, Mesh.cpp:64-89; Rectangle.cpp:1255-1613)
// in one kernel, bit-identical to the sub-step kernels and the CPU oracle.
//
// Design (register streaming, no LDS):
//  * A wavefront owns 64 consecutive p-cells, lane l <-> j = jw - 3 + l.  Every
//    quantity of the stage depends on p-neighbours at most 3 cells away, so
//    lanes 3..60 produce exact results (58 outputs per wavefront) and p
//    neighbours are read from other lanes with cross-lane shuffles.
//  * The wavefront marches along x over a strip of TI output rows.  At march
//    row r it loads f_s(r) and f^n(r) (one coalesced 512-byte row each) and
//    advances a software pipeline held in registers:
//        x-face r-1 : WENO value and high-order flux FxH_s
//        column r   : WENO value; column r-1: high-order flux FpH_s
//        x-face r, column r: low-order fluxes of stage 0 (recomputed from f^n
//                     and the stage-0 fields instead of stored)
//        cell r-1   : low-order predictor f2
//        x-face r-1, column r-1: antidiffusive fluxes FxDS/FpDS
//        cell r-2   : Zalesak R+-
//        cell r-3   : limiter C on its faces and the corrected value f_{s+1}
//  * Store-flux design: FxH_s/FpH_s are written once (owned faces) for the
//    later stages' RK combinations instead of recomputing the WENO fluxes of
//    every earlier stage from stored stage values.  Stage 5 needs the earlier
//    fluxes only through its own combination, so stage 4 writes, in place of
//    FxH_4/FpH_4, stage 5's partial sum b_0 F_0 + ... + b_4 F_4 (left to
//    right, the order stage 5 would use), and stage 5 reads that one plane per
//    direction instead of five: the same bits, 64 B/cell less to read and
//    fewer registers held at stage 5.
// Every quantity is computed only inside the reference's loop range and is 0
// outside it, as the reference's never-written array entries are.
//
// Mixed-precision variant (MP): the divisions of the stage (the WENO weights
// of both face reconstructions, and the Zalesak ratios R+-) run in FP32 on f
// scaled by an exact power of two (f * 2^-k with 2^k near the patch's max |f|);
// the candidate face values, the fluxes, the RK combination and every
// accumulation into f stay FP64.  A face value is fR + w (fL - fR) with the
// FP32 weight w, so its error is of order 1e-7 |fL - fR|, not 1e-7 |f|.

#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "vx_dev.h"
#include "vx_internal.h"

using namespace vxi;
using namespace vxd;

namespace {

constexpr int HALO = 3;
constexpr int OUTW = 64 - 2 * HALO;  // p-cells produced per wavefront

struct StageArgs {
    vx_patch p;
    const double *fin;  // f_s (stage input), ghost ring zero
    double *fout;       // f_{s+1}; only interior cells are written
    const double *E, *a2;    // fields of this stage (finest grid)
    const double *E0, *a20;  // fields of stage 0 (for the low-order fluxes)
    int n_f;
    int step, ns;
    double a[6], aSum;
    double b5[6];  // stage 5's coefficients times dt (stage 4 forms its partial sum)
    int TI;       // output rows per strip
    int w4;       // 1: the kernel form held to four waves per SIMD (k_stage_w4)
    int nstrip;   // strips in x
    int nbx;      // tiles along p
    double sinv;  // MP: 2^-k, the scale of the FP32 copies of f
    float eps32;  // MP: the WENO epsilon (1e-10) in the scaled units
};

__device__ __forceinline__ double ldp(const double *pl, int pitch, int nx, int np, int i, int j) {
    return (i >= -2 && i < nx + 2 && j >= -2 && j < np + 2) ? pl[pidx(pitch, i, j)] : 0.0;
}

// Row i of a plane at the clamped column jc, with i clamped into the plane:
// no branch around the load, so that the loads in flight are counted one by
// one and a row waits only for its own; the caller masks the value (0 outside
// the plane) where it uses it, as ldp would have.
__device__ __forceinline__ double ldr(const double *pl, int pitch, int nx, int i, int jc) {
    const int ic = i < -2 ? -2 : (i > nx + 1 ? nx + 1 : i);
    return pl[pidx(pitch, ic, jc)];
}

// q^2 A^2 at x-face i (Rectangle.cpp:1280, 1296-1297; EMSolver.cpp:133-135)
__device__ __forceinline__ double face_as(const StageArgs &A, const double *a2, int i) {
    int fi = A.p.rtb * (i + A.p.x_pos);
    int c = fi < 0 ? 0 : (fi > A.n_f ? A.n_f : fi);
    return A.p.q * A.p.q * a2[c];
}

// q * Rectangle::GetEfield(i) (Rectangle.cpp:1055-1067, 1298)
__device__ __forceinline__ double col_qE(const StageArgs &A, const double *E, int i) {
    return A.p.q * patch_E(E, A.n_f, A.p.rtb, A.p.x_pos, i);
}

// XCD-aware tile order (vxd::xcd_order): p-neighbouring tiles march over the
// same rows on the same XCD and share their overlapping halos in its L2.
__device__ __forceinline__ int xcd_tile() { return xcd_order((int)blockIdx.x, (int)gridDim.x); }

// A strip of one wavefront: TI output rows [i0, i1) of the 58 p-cells from jw.
// IN (interior): the strip and its p-column lie so far from the patch edges
// that every reference loop-range test below holds for every lane and every
// march row (interior_strip states the margins), so IN only drops tests that
// are true, and the IN and general forms give the same bits.
template <int STEP, bool MP, bool IN>
__device__ __forceinline__ void stage_strip(const StageArgs &A, const int jw, const int i0) {
    const vx_patch &P = A.p;
    const int nx = P.n_x, np = P.n_p, pitch = P.pitch;
    const int lane = threadIdx.x & 63;
    const int i1 = min(i0 + A.TI, nx);       // output rows [i0, i1)
    const bool last_strip = i0 + A.TI >= nx;
    const int j = jw - HALO + lane;
    const bool mid_lane = lane >= HALO && lane < 64 - HALO;
    const bool out_lane = mid_lane && (IN || j < np);
    // faces this wavefront stores (each valid face has exactly one owner)
    const bool own_row_x = IN ? mid_lane
                              : (mid_lane && j < np) || (j == np && lane >= HALO && lane < 65 - HALO) ||
                                    (jw == 0 && j == -1);
    const bool own_row_p = IN ? mid_lane : (mid_lane && j < np) || (j == np && lane >= HALO && lane < 65 - HALO);
    constexpr bool store_flux = STEP < 5;

    // reference loop ranges in j (Rectangle.cpp:1279-1394) for this lane
    const bool jr_ex = IN || (j >= -1 && j < np + 1);   // ex, fx, FxH, FxL rows
    const bool jr_ep = IN || (j >= -1 && j < np + 2);   // ep
    const bool jr_fp = IN || (j >= 0 && j < np + 1);    // fp, FpH, FpL p-faces
    const bool jr_R = IN || (j >= -1 && j < np + 1);    // R
    // update loop ranges (walls on all sides: xm = 1, xp = n_x, pm = 1, pp = n_p)
    const int xm = P.left ? 1 : 0, xp = P.right ? nx : nx + 1;
    const int pm = P.down ? 1 : 0, pp = P.up ? np : np + 1;

    const double mc2 = P.mc2inv;
    const double dpcc = P.dp_inv * P.cc;
    const double ccdx = P.cc * P.dx_inv;
    const double w3 = 1 / 48.0;
    const double pj = momentum(P.pmin, P.dp, j, P.p_pos);
    const double pj2 = pj * pj;

    const double *fin = A.fin, *f0p = P.f[0];
    const int r0 = i0 - HALO;

    // pipeline state (names: _mK = row/face r-K)
    double fs_m3 = 0, fs_m2 = 0, fs_m1 = 0, fs_0 = 0;
    double f0_m3 = 0, f0_m2 = 0, f0_m1 = 0, f0_0 = 0;
    double ep_m2 = 0, ep_m1 = 0, ep_0 = 0, fp_m2 = 0, fp_m1 = 0, fp_0 = 0;
    double FxL_m1 = 0, FxL_0 = 0, FpL_m1 = 0, FpL_0 = 0;
    double f2_m3 = 0, f2_m2 = 0, f2_m1 = 0;
    double FxD_m3 = 0, FxD_m2 = 0, FxD_m1 = 0;
    double FpD_m3 = 0, FpD_m2 = 0, FpD_m1 = 0;
    double Rp_m4 = 0, Rp_m3 = 0, Rp_m2 = 0, Rm_m4 = 0, Rm_m3 = 0, Rm_m2 = 0;
    float gs_m3 = 0, gs_m2 = 0, gs_m1 = 0, gs_0 = 0;  // MP: FP32 copies of f_s rows
    const double sinv = A.sinv;
    const float eps32 = A.eps32;

    // Gamma(p_j, q^2 A^2 at x-face i), current fields and stage-0 fields
    double as_m1 = face_as(A, A.a2, r0 - 1), as_0 = face_as(A, A.a2, r0);
    double G_m1 = gamma_(pj, as_m1, mc2);
    double G_0 = (as_0 == as_m1) ? G_m1 : gamma_(pj, as_0, mc2);
    double as0_0 = face_as(A, A.a20, r0);
    double G0_0 = (STEP == 0 || as0_0 == as_0) ? G_0 : gamma_(pj, as0_0, mc2);

    // software prefetch: rows r and r+1 are in flight before iteration r, the
    // stored fluxes of x-face / column r-1 one iteration ahead
    const bool in_j = IN || (j >= -2 && j < np + 2);
    const int jc = IN ? j : (j < -2 ? -2 : (j > np + 1 ? np + 1 : j));
    auto row_ok = [&](int i) { return IN || (i >= -2 && i < nx + 2 && in_j); };
    double pfs_a = ldr(fin, pitch, nx, r0, jc), pf0_a = ldr(f0p, pitch, nx, r0, jc);
    double pfs_b = ldr(fin, pitch, nx, r0 + 1, jc), pf0_b = ldr(f0p, pitch, nx, r0 + 1, jc);
    // stored flux planes read by this stage; stage 5 reads only plane 4, which
    // holds its partial sum
    constexpr int NH = STEP == 5 ? 1 : (STEP > 0 ? STEP : 1);
    double hx[NH], hp[NH];
    bool hvx = false, hvp = false;  // the loaded face lies in the FxH / FpH ranges
    auto load_h = [&](int face) {
        hvx = IN || (face >= 0 && face < nx + 1 && jr_ex);
        hvp = IN || (face >= -1 && face < nx + 1 && jr_fp);
#pragma unroll
        for (int k = 0; k < NH; k++) {
            const int plane = STEP == 5 ? 4 : k;
            hx[k] = (k < STEP) ? ldr(P.FxH[plane], pitch, nx, face, jc) : 0.0;
            hp[k] = (k < STEP) ? ldr(P.FpH[plane], pitch, nx, face, jc) : 0.0;
        }
    };
    load_h(r0 - 1);

    // The RK combination sum_k a_k FxH_k is evaluated left to right; the terms
    // of the earlier stages are folded into Sx/Sp as soon as their prefetched
    // values are consumed, which frees their registers a whole iteration early.
    // Stage 4 folds the same values into stage 5's partial sum Qx/Qp.
    double Sx = 0.0, Sp = 0.0, Qx = 0.0, Qp = 0.0;

    // One march row; each quantity is masked to its reference loop range.
    auto row = [&](int r) {
        // ---- ingest row r
        fs_0 = row_ok(r) ? pfs_a : 0.0;
        f0_0 = row_ok(r) ? pf0_a : 0.0;
        if (MP) gs_0 = (float)(fs_0 * sinv);
        pfs_a = pfs_b;
        pf0_a = pf0_b;
        pfs_b = ldr(fin, pitch, nx, r + 2, jc);
        pf0_b = ldr(f0p, pitch, nx, r + 2, jc);
        if (STEP == 5) {
            // b_0 F_0 + ... + b_4 F_4 as stage 4 formed it; outside the face
            // ranges the masked sum of the five terms is +0
            Sx = hvx ? hx[0] : 0.0;
            Sp = hvp ? hp[0] : 0.0;
        } else if (STEP > 0) {
            Sx = A.a[0] * (hvx ? hx[0] : 0.0);
            Sp = A.a[0] * (hvp ? hp[0] : 0.0);
#pragma unroll
            for (int k = 1; k < NH; k++) {
                Sx = Sx + A.a[k] * (hvx ? hx[k] : 0.0);
                Sp = Sp + A.a[k] * (hvp ? hp[k] : 0.0);
            }
            if (STEP == 4) {
                Qx = A.b5[0] * (hvx ? hx[0] : 0.0);
                Qp = A.b5[0] * (hvp ? hp[0] : 0.0);
#pragma unroll
                for (int k = 1; k < NH; k++) {
                    Qx = Qx + A.b5[k] * (hvx ? hx[k] : 0.0);
                    Qp = Qp + A.b5[k] * (hvp ? hp[k] : 0.0);
                }
            }
        }
        load_h(r);
        double as_p1 = face_as(A, A.a2, r + 1);
        double G_p1 = (as_p1 == as_0) ? G_0 : gamma_(pj, as_p1, mc2);
        double as0_p1 = face_as(A, A.a20, r + 1);
        double G0_p1 = (STEP == 0 || as0_p1 == as_p1) ? G_p1 : gamma_(pj, as0_p1, mc2);

        // ---- x-face r-1: ex, fx (Rectangle.cpp:1278-1293), FxH_s (1313-1334)
        const int ix = r - 1;
        const bool vFx = IN || (ix >= 0 && ix < nx + 1 && jr_ex);
        double Gd = shdn1(G_m1);  // Gamma at p-face j+1
        double ex = (IN || (ix >= -1 && ix < nx + 2 && jr_ex)) ? dpcc * (Gd - G_m1) : 0.0;
        double fx = 0.0;
        if (vFx) {
            if (MP) fx = weno_mp(fs_m3, fs_m2, fs_m1, fs_0, gs_m3, gs_m2, gs_m1, gs_0, eps32, ex > 0.0);
            else fx = weno(fs_m3, fs_m2, fs_m1, fs_0, ex > 0.0);
        }
        double ex_u = shdn1(ex), ex_d = shup1(ex);
        double fx_u = shdn1(fx), fx_d = shup1(fx);
        double FxH = vFx ? P.dx_inv * (fx * ex + w3 * (fx_u - fx_d) * (ex_u - ex_d)) : 0.0;

        // ---- column r: ep, fp (Rectangle.cpp:1294-1312)
        const bool cr = IN || (r >= -1 && r < nx + 1);
        double qE = cr ? col_qE(A, A.E, r) : 0.0;
        ep_0 = (cr && jr_ep) ? qE - ccdx * (G_p1 - G_0) : 0.0;
        double s1 = shup1(fs_0), s2 = shup1(s1), s3 = shdn1(fs_0);
        fp_0 = 0.0;
        if (MP) {
            float t1 = shup1f(gs_0), t2 = shup1f(t1), t3 = shdn1f(gs_0);
            if (cr && jr_fp) fp_0 = weno_mp(s2, s1, fs_0, s3, t2, t1, gs_0, t3, eps32, ep_0 > 0.0);
        } else if (cr && jr_fp) {
            fp_0 = weno(s2, s1, fs_0, s3, ep_0 > 0.0);
        }
        // ---- column r-1: FpH_s (Rectangle.cpp:1353-1375)
        const bool vFp = IN || (ix >= -1 && ix < nx + 1 && jr_fp);
        double FpH = vFp ? P.dp_inv * (fp_m1 * ep_m1 + w3 * (fp_0 - fp_m2) * (ep_0 - ep_m2)) : 0.0;
        if (store_flux) {
            const bool own_x = ix >= i0 && (ix < i0 + A.TI || last_strip);
            const bool own_c = (ix >= i0 || i0 == 0) && (ix < i0 + A.TI || last_strip);
            // read again only by the later stages: streamed past the caches;
            // stage 4 stores stage 5's partial sum instead of its fluxes
            const double hx_out = STEP == 4 ? Qx + A.b5[4] * FxH : FxH;
            const double hp_out = STEP == 4 ? Qp + A.b5[4] * FpH : FpH;
            if (vFx && own_x && own_row_x) __builtin_nontemporal_store(hx_out, &P.FxH[STEP][pidx(pitch, ix, j)]);
            if (vFp && own_c && own_row_p) __builtin_nontemporal_store(hp_out, &P.FpH[STEP][pidx(pitch, ix, j)]);
        }

        // ---- low-order fluxes of stage 0 at x-face r and column r
        // (Rectangle.cpp:1335-1352, 1376-1394 evaluated with the stage-0 fields)
        double G0d = shdn1(G0_0);
        double ex0 = dpcc * (G0d - G0_0);
        FxL_0 = (IN || (r >= 0 && r < nx + 1 && jr_ex)) ? P.dx_inv * ((ex0 > 0.0 ? f0_m1 : f0_0) * ex0) : 0.0;
        double qE0 = cr ? col_qE(A, A.E0, r) : 0.0;
        double ep0 = qE0 - ccdx * (G0_p1 - G0_0);
        double f0_jm1 = shup1(f0_0);
        FpL_0 = (cr && jr_fp) ? P.dp_inv * ((ep0 > 0.0 ? f0_jm1 : f0_0) * ep0) : 0.0;

        // ---- cell r-1: low-order predictor (Rectangle.cpp:1518-1534, gather form)
        {
            const int i = r - 1;
            double FpL_up = shdn1(FpL_m1);  // FpL0 at (i, j+1)
            double t = 0.0;
            if (IN || (i >= 0 && i < nx && j >= 0 && j < np)) {
                t = f0_m1;
                const bool in0 = IN || (i >= xm && i < xp && j >= pm && j < pp);
                if (in0) {
                    t += A.aSum * FxL_m1;
                    t += A.aSum * FpL_m1;
                }
                if (IN || (i >= xm && i < xp && j + 1 >= pm && j + 1 < pp)) t -= A.aSum * FpL_up;
                if (IN || (i + 1 >= xm && i + 1 < xp && j >= pm && j < pp)) t -= A.aSum * FxL_0;
            }
            f2_m1 = t;  // ghost cells: zero fill of the walls (BoundaryCondition.cpp:6-8)
        }

        // ---- x-face r-1 and column r-1: RK flux combination (Rectangle.cpp:1396-1517)
        {
            double xLS = A.aSum * FxL_m1, pLS = A.aSum * FpL_m1;
            double sx = STEP == 0 ? A.a[0] * FxH : Sx + A.a[STEP] * FxH;
            double sp = STEP == 0 ? A.a[0] * FpH : Sp + A.a[STEP] * FpH;
            FxD_m1 = sx - xLS;
            FpD_m1 = sp - pLS;
        }

        // ---- cell r-2: Zalesak R+- (Rectangle.cpp:1538-1579)
        {
            const int i = r - 2;
            double FpDu = shdn1(FpD_m2);  // FpDS at (i, j+1)
            double f0u = shdn1(f0_m2), f0d = shup1(f0_m2);
            double f2u = shdn1(f2_m2), f2d = shup1(f2_m2);
            double rp = 0.0, rm = 0.0;
            if (IN || (i >= -1 && i < nx + 1 && jr_R)) {
                // fmax/fmin (one v_max_f64/v_min_f64 each) give the reference's
                // a > b ? a : b bits here: f is never -0 (every update adds to a
                // +0 or positive start), R and Q are never -0, nothing is NaN,
                // and a signed zero among the P terms reaches P only when P is
                // zero, where R is 0 either way
                double Pp = fmax(0.0, FxD_m2) - fmin(0.0, FxD_m1) + fmax(0.0, FpD_m2) - fmin(0.0, FpDu);
                double Pm = fmax(0.0, FxD_m1) - fmin(0.0, FxD_m2) + fmax(0.0, FpDu) - fmin(0.0, FpD_m2);
                double wMax = hwmax(hwmax(f0_m2, f2_m2), hwmax(hwmax(f0_m1, f2_m1),
                              hwmax(hwmax(f0_m3, f2_m3), hwmax(hwmax(f0u, f2u), hwmax(f0d, f2d)))));
                double wMin = hwmin(hwmin(f0_m2, f2_m2), hwmin(hwmin(f0_m1, f2_m1),
                              hwmin(hwmin(f0_m3, f2_m3), hwmin(hwmin(f0u, f2u), hwmin(f0d, f2d)))));
                double Qm = -wMin + f2_m2;
                double Qp = wMax - f2_m2;
                if (MP) {
                    rp = Pp > 0.0 ? zalesak_mp(Qp, Pp, sinv) : 0.0;
                    rm = Pm > 0.0 ? zalesak_mp(Qm, Pm, sinv) : 0.0;
                } else {
                    rp = Pp > 0.0 ? fmin(1.0, Qp / Pp) : 0.0;
                    rm = Pm > 0.0 ? fmin(1.0, Qm / Pm) : 0.0;
                }
            }
            Rp_m2 = rp;
            Rm_m2 = rm;
        }

        // ---- cell r-3: limiter C on its faces (Rectangle.cpp:1581-1594) and the
        // corrected value (Rectangle.cpp:1596-1612, gather form)
        {
            const int i = r - 3;
            double Rpd = shup1(Rp_m3), Rmd = shup1(Rm_m3);  // R at (i, j-1)
            const bool jint = IN || (j >= 0 && j < np);
            double cxl = (IN || (i >= 1 && i < nx && jint))
                             ? (FxD_m3 > 0.0 ? hwmin(Rp_m3, Rm_m4) : hwmin(Rp_m4, Rm_m3)) : 1.0;
            double cxr = (IN || (i + 1 >= 1 && i + 1 < nx && jint))
                             ? (FxD_m2 > 0.0 ? hwmin(Rp_m2, Rm_m3) : hwmin(Rp_m3, Rm_m2)) : 1.0;
            double cpl = (IN || (i >= 1 && i < nx && jint))
                             ? (FpD_m3 > 0.0 ? hwmin(Rp_m3, Rmd) : hwmin(Rpd, Rm_m3)) : 1.0;
            double cpu = shdn1(cpl);      // C at (i, j+1), computed by the next lane
            double FpDu = shdn1(FpD_m3);  // FpDS at (i, j+1)
            if (out_lane && i >= i0 && i < i1) {
                double t = f2_m3;
                if (IN || (i >= xm && i < xp && j >= pm && j < pp)) {
                    t += cxl * FxD_m3;
                    t += cpl * FpD_m3;
                }
                if (IN || (i >= xm && i < xp && j + 1 >= pm && j + 1 < pp)) t -= cpu * FpDu;
                if (IN || (i + 1 >= xm && i + 1 < xp && j >= pm && j < pp)) t -= cxr * FxD_m2;
                A.fout[pidx(pitch, i, j)] = t;
            }
        }

        // ---- shift the pipeline by one row
        fs_m3 = fs_m2; fs_m2 = fs_m1; fs_m1 = fs_0;
        gs_m3 = gs_m2; gs_m2 = gs_m1; gs_m1 = gs_0;
        f0_m3 = f0_m2; f0_m2 = f0_m1; f0_m1 = f0_0;
        as_0 = as_p1; G_m1 = G_0; G_0 = G_p1; G0_0 = G0_p1;
        ep_m2 = ep_m1; ep_m1 = ep_0; fp_m2 = fp_m1; fp_m1 = fp_0;
        FxL_m1 = FxL_0; FpL_m1 = FpL_0;
        f2_m3 = f2_m2; f2_m2 = f2_m1;
        FxD_m3 = FxD_m2; FxD_m2 = FxD_m1;
        FpD_m3 = FpD_m2; FpD_m2 = FpD_m1;
        Rp_m4 = Rp_m3; Rp_m3 = Rp_m2; Rm_m4 = Rm_m3; Rm_m3 = Rm_m2;
    };

    for (int r = r0; r <= i1 + HALO - 1; r++) row(r);
}

// Whether a strip takes the IN form.  Lanes hold j in [jw - 3, jw + 60] and
// the march rows are r in [i0 - 3, i1 + 2] with i1 = i0 + TI; the tightest
// tests above are the updates of cells (r - 1, j) and (r - 3, j) with their
// +1 neighbours inside [1, n - 1) (xm, pm <= 1, xp >= n_x, pp >= n_p), which
// need 1 <= j <= n_p - 2 and 4 <= r <= n_x - 1, and the stored-flux faces
// r0 - 1 >= 0.  The margins below are wider than that.
__device__ __forceinline__ bool interior_strip(const StageArgs &A, int jw, int i0) {
    return jw >= 8 && jw + 64 <= A.p.n_p && i0 >= 8 && i0 + A.TI + 8 <= A.p.n_x;
}

// One workgroup's tile (4 strips of one 64-lane p-column) of a stage.
template <int STEP, bool MP>
__device__ __forceinline__ void stage_tile(const StageArgs &A, const int tile) {
    // the wavefront index is uniform: row loop, row ranges and field reads stay scalar
    const int bx = tile % A.nbx, by = tile / A.nbx;
    const int strip = by * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (strip >= A.nstrip) return;  // uniform per wavefront
    const int jw = bx * OUTW, i0 = strip * A.TI;
    if (interior_strip(A, jw, i0)) stage_strip<STEP, MP, true>(A, jw, i0);
    else stage_strip<STEP, MP, false>(A, jw, i0);
}

// Compiled for at least three resident waves per SIMD (at most 168 VGPRs):
// the FP64 stages fit as they are; the mixed-precision stage 4 (173 otherwise)
// keeps a few values in scratch instead of running two waves per SIMD.
template <int STEP, bool MP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_stage(StageArgs A) {
    stage_tile<STEP, MP>(A, xcd_tile());
}
// The same stage compiled for four resident waves per SIMD (at most 128
// VGPRs; some stages then keep a few loop-invariant values in scratch): more
// waves to hide latency against a few more loads.  The strip-height timing
// picks between the two forms per patch shape and stage; both give the same bits.
template <int STEP, bool MP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_stage_w4(StageArgs A) {
    stage_tile<STEP, MP>(A, xcd_tile());
}

// Several patches (of independent runs) at the same RK stage in one launch:
// workgroups [wg_start[k], wg_start[k+1]) of the XCD-ordered tile sequence
// belong to item k.  Held to three waves per SIMD like k_stage (stages 3 and
// 4 would take 185-206 VGPRs, two waves).
template <int STEP, bool MP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_stage_batch(const StageArgs *items, const int *wg_start, int n) {
    const int tile = xcd_tile();
    int lo = 0, hi = n - 1;  // the last k with wg_start[k] <= tile (uniform)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wg_start[mid] <= tile) lo = mid;
        else hi = mid - 1;
    }
    stage_tile<STEP, MP>(items[lo], tile - wg_start[lo]);
}

template <bool MP>
void launch_stage_t(vx_ctx *ctx, const StageArgs &A, dim3 grid) {
#define VX_STAGE(S)                                                        \
    if (A.w4) k_stage_w4<S, MP><<<grid, 256, 0, ctx->stream>>>(A);         \
    else k_stage<S, MP><<<grid, 256, 0, ctx->stream>>>(A);
    switch (A.step) {
        case 0: VX_STAGE(0) break;
        case 1: VX_STAGE(1) break;
        case 2: VX_STAGE(2) break;
        case 3: VX_STAGE(3) break;
        case 4: VX_STAGE(4) break;
        default: VX_STAGE(5) break;
    }
#undef VX_STAGE
}

void launch_stage(vx_ctx *ctx, StageArgs &A, bool mp) {
    A.nstrip = (A.p.n_x + A.TI - 1) / A.TI;
    dim3 grid((unsigned)(A.nbx * ((A.nstrip + 3) / 4)));
    if (mp) launch_stage_t<true>(ctx, A, grid);
    else launch_stage_t<false>(ctx, A, grid);
}

// Rows per strip (the tile height) of a patch shape and RK stage (the later
// stages hold more registers and read more planes).  A strip costs its rows
// plus 6 pipeline warm-up rows, and a wavefront walks them one after another,
// so small patches want short strips (more wavefronts) and large ones long
// strips (less warm-up).  The first call for a shape times the candidates on
// the call's own launch, which is idempotent (it reads f_s, f^n and the fluxes
// of earlier stages and writes only f_(s+1) and this stage's fluxes, the same
// values for every strip height), and keeps the fastest (a shorter strip must
// beat the taller ones by more than 3 %).  While a step graph
// is being recorded the choice falls back to the shortest strip that still
// gives 2048 wavefronts.
//
// Each strip height is timed in both kernel forms (k_stage: three waves per
// SIMD for the stages that need more than 128 VGPRs; k_stage_w4: four), the
// four-wave form winning only when more than 3 % faster; a choice is kept as
// rows + W4_FORM for the four-wave form.
// The choices are kept per device for the whole process (every context on the
// device shares them: a serving loop's runs of one shape time it once).
struct TileKey {
    int device, nx, np, mp, step;
    bool operator<(const TileKey &o) const {
        return std::tie(device, nx, np, mp, step) < std::tie(o.device, o.nx, o.np, o.mp, o.step);
    }
};
constexpr int W4_FORM = 4096;
std::mutex g_tile_mu;
std::map<TileKey, int> g_tiles;

int tile_lookup(int device, int nx, int np, int mp, int step) {
    std::lock_guard<std::mutex> g(g_tile_mu);
    auto it = g_tiles.find(TileKey{device, nx, np, mp, step});
    return it == g_tiles.end() ? 0 : it->second;
}

void tile_store(vx_ctx *ctx, int nx, int np, int ti, int mp, int step) {
    ctx->ti_cache.push_back({nx, np, ti, mp, step});
    std::lock_guard<std::mutex> g(g_tile_mu);
    g_tiles[TileKey{ctx->device, nx, np, mp, step}] = ti;
}

// Workgroups of k_stage<step, mp> resident per compute unit (0 if unknown).
int stage_blocks_per_cu(int step, bool mp) {
    int n = 0;
    hipError_t e = hipSuccess;
#define VX_OCC(S, M) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage<S, M>, 256, 0)
    if (mp) {
        switch (step) {
            case 0: VX_OCC(0, true); break;
            case 1: VX_OCC(1, true); break;
            case 2: VX_OCC(2, true); break;
            case 3: VX_OCC(3, true); break;
            case 4: VX_OCC(4, true); break;
            default: VX_OCC(5, true); break;
        }
    } else {
        switch (step) {
            case 0: VX_OCC(0, false); break;
            case 1: VX_OCC(1, false); break;
            case 2: VX_OCC(2, false); break;
            case 3: VX_OCC(3, false); break;
            case 4: VX_OCC(4, false); break;
            default: VX_OCC(5, false); break;
        }
    }
#undef VX_OCC
    return e == hipSuccess ? n : 0;
}

int tile_rows(vx_ctx *ctx, const StageArgs &A0, bool mp) {
    const int nx = A0.p.n_x, np = A0.p.n_p;
    for (auto &c : ctx->ti_cache)
        if (c[0] == nx && c[1] == np && c[3] == (int)mp && c[4] == A0.step) return c[2];
    static const int cand[] = {64, 32, 16, 8, 4, 2};
    // VX_TILE_ROWS pins the choice for this context (one value, or six
    // comma-separated values, one per RK stage; rows + 4096 for the four-wave
    // form), e.g. to profile the configuration a timed run chose
    if (const char *env = std::getenv("VX_TILE_ROWS")) {
        int v[6], n = 0;
        for (const char *q = env; *q && n < 6;) {
            v[n++] = std::atoi(q);
            while (*q && *q != ',') q++;
            if (*q == ',') q++;
        }
        const int ti = n == 6 ? v[A0.step] : (n >= 1 ? v[0] : 0);
        if ((ti >= 1 && ti <= 1024) || (ti > W4_FORM && ti <= W4_FORM + 1024)) {
            ctx->ti_cache.push_back({nx, np, ti, (int)mp, A0.step});  // this context only
            return ti;
        }
    }
    if (const int known = tile_lookup(ctx->device, nx, np, (int)mp, A0.step)) {
        ctx->ti_cache.push_back({nx, np, known, (int)mp, A0.step});
        return known;
    }
    int guess = 2;
    for (int ti : cand)
        if ((long)A0.nbx * ((nx + ti - 1) / ti) >= 2048) {
            guess = ti;
            break;
        }
    if (ctx->capturing) return guess;
    // Besides the powers of two, the strip heights at which the launch's
    // workgroups just fill k rounds of the device's resident slots (k = 1..4):
    // the tallest strips (least warm-up) without a part-empty last round.
    std::vector<int> cands(std::begin(cand), std::end(cand));
    const int slots = ctx->n_cu * stage_blocks_per_cu(A0.step, mp);
    if (slots > 0)
        for (int k = 1; k <= 4; k++) {
            const long wg_max = (long)k * slots;
            // smallest strip height whose ceil(strips / 4) * nbx workgroups fit in wg_max
            for (int ti = 1; ti <= 1024 && ti <= nx; ti++) {
                const long strips = (nx + ti - 1) / ti;
                if ((long)A0.nbx * ((strips + 3) / 4) <= wg_max) {
                    if (ti > 2) cands.push_back(ti);
                    break;
                }
            }
        }
    std::sort(cands.begin(), cands.end(), std::greater<int>());
    cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = guess;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        bool failed = false;
        // one launch of (strip height, form) timed on its own
        auto timed = [&](int code) {
            StageArgs A = A0;
            A.TI = code % W4_FORM;
            A.w4 = code / W4_FORM;
            float t = 1e30f;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) failed = true;
            if (!failed) launch_stage(ctx, A, mp);
            if (!failed && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&t, e0, e1) != hipSuccess))
                failed = true;
            return t;
        };
        // Candidates run from tall to short strips, the three-wave form first.
        // One warm-up launch, then one timed launch of each candidate; the
        // three fastest are timed twice more and keep their fastest launch
        // (the minimum is steadier than a sum against the device's noise).
        // Among those, in candidate order, a later one must be more than 3 %
        // faster to win (ties go to fewer warm-up rows and the three-wave form).
        std::vector<std::pair<int, float>> res;  // (code, ms) in candidate order
        {
            StageArgs A = A0;
            A.TI = guess;
            A.w4 = 0;
            launch_stage(ctx, A, mp);  // warm
        }
        for (int ti : cands) {
            if (ti > 2 && ti >= 2 * nx) continue;  // a single strip either way
            for (int w4 = 0; w4 < 2 && !failed; w4++) {
                const int code = ti + (w4 ? W4_FORM : 0);
                const float t = timed(code);
                if (!failed) res.push_back({code, t});
            }
            if (failed) break;
        }
        if (!failed && !res.empty()) {
            std::vector<int> order(res.size());
            for (size_t k = 0; k < res.size(); k++) order[k] = (int)k;
            std::sort(order.begin(), order.end(), [&](int a, int b) { return res[a].second < res[b].second; });
            const size_t nfin = std::min<size_t>(3, order.size());
            std::vector<int> fin(order.begin(), order.begin() + nfin);
            for (int k : fin)
                for (int rep = 0; rep < 2 && !failed; rep++) res[k].second = std::min(res[k].second, timed(res[k].first));
            if (!failed) {
                std::sort(fin.begin(), fin.end());  // candidate order
                float best_ms = 1e30f;
                for (int k : fin)
                    if (res[k].second < 0.97f * best_ms) {
                        best_ms = res[k].second;
                        best = res[k].first;
                    }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    tile_store(ctx, nx, np, best, (int)mp, A0.step);
    return best;
}

}  // namespace

extern "C" {

int vx_fct_stage_fused_available(void) { return 1; }

}  // extern "C"

// Checks and kernel arguments of one patch's fused stage.
static int stage_args(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl,
                      const double *fin, double *fout, bool mp, double f_scale, StageArgs &A) {
    int rc = validate_patch(ctx, p);
    if (rc) return rc;
    if (step < 0 || step > 5 || !fl || !fl->E0 || !fl->a20)
        return fail(ctx, VX_EINVAL, "stage_fused: bad step %d or missing stage-0 fields", step);
    if ((rc = validate_span(ctx, p, fl->n_f))) return rc;
    if (!(p->left && p->right && p->down && p->up))
        return fail(ctx, VX_EINVAL, "stage_fused: patch must touch all four domain edges");
    if (fin == p->f[0] && step != 0)
        return fail(ctx, VX_EINVAL, "stage_fused: input aliases f^n after stage 0");
    if (fout == fin || fout == p->f[0])
        return fail(ctx, VX_EINVAL, "stage_fused: output must not alias the inputs");
    A.p = *p;
    A.fin = fin;
    A.fout = fout;
    A.E = fl->E;
    A.a2 = fl->a2;
    A.E0 = fl->E0;
    A.a20 = fl->a20;
    A.n_f = fl->n_f;
    A.step = step;
    A.ns = step + 1;
    for (int k = 0; k < 6; k++) A.a[k] = 0.0;
    for (int k = 0; k < A.ns; k++) A.a[k] = vx::rk_a[step][k] * dt;
    for (int k = 0; k < 6; k++) A.b5[k] = vx::rk_a[5][k] * dt;  // as stage 5 forms its a[k]
    A.aSum = A.a[0];
    for (int k = 1; k < A.ns; k++) A.aSum = A.aSum + A.a[k];
    A.nbx = (p->n_p + OUTW - 1) / OUTW;
    A.sinv = 1.0;
    A.eps32 = 0.0f;
    if (mp) {
        if (!(f_scale > 0.0) || !std::isfinite(f_scale))
            return fail(ctx, VX_EINVAL, "stage_fused_mp: f scale must be positive and finite");
        A.sinv = std::ldexp(1.0, -std::ilogb(f_scale));
        A.eps32 = (float)(1.0e-10 * A.sinv * A.sinv);
    }
    A.w4 = 0;
    const int choice = tile_rows(ctx, A, mp);
    A.TI = choice % W4_FORM;
    A.w4 = choice / W4_FORM;
    A.nstrip = (p->n_x + A.TI - 1) / A.TI;
    return VX_OK;
}

static int stage_fused(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl,
                       const double *fin, double *fout, bool mp, double f_scale) {
    StageArgs A;
    int rc = stage_args(ctx, p, step, dt, fl, fin, fout, mp, f_scale, A);
    if (rc) return rc;
    LaunchTimer lt(ctx, "k_stage");
    launch_stage(ctx, A, mp);
    return check_launch(ctx, "k_stage");
}

// Staging of batched launch arguments: device slots with pinned host mirrors;
// a slot is rewritten only after the copy that last used it has completed.
constexpr int BATCH_MAX = 64, BATCH_SLOTS = 8;
constexpr size_t BATCH_SLOT_BYTES = sizeof(StageArgs) * BATCH_MAX + sizeof(int) * (BATCH_MAX + 1) + 256;

struct BatchRing {
    char *dev = nullptr, *host = nullptr;
    hipEvent_t ev[BATCH_SLOTS] = {};
    bool pending[BATCH_SLOTS] = {};
    int next = 0;
};

static void batch_ring_free(void *r) {
    BatchRing *b = (BatchRing *)r;
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->dev) (void)hipFree(b->dev);
    if (b->host) (void)hipHostFree(b->host);
    delete b;
}

static int batch_launch(vx_ctx *ctx, const std::vector<StageArgs> &A, int step, bool mp) {
    if (A.empty()) return VX_OK;
    if (ctx->capturing) return fail(ctx, VX_EINVAL, "stage batch while a step graph is recorded");
    if (!ctx->batch_ring) {
        BatchRing *b = new BatchRing;
        ctx->batch_ring = b;
        ctx->batch_ring_free = batch_ring_free;
        int rc = check_hip(ctx, hipMalloc((void **)&b->dev, BATCH_SLOT_BYTES * BATCH_SLOTS), "batch args");
        if (!rc) rc = check_hip(ctx, hipHostMalloc((void **)&b->host, BATCH_SLOT_BYTES * BATCH_SLOTS), "batch staging");
        for (auto &e : b->ev)
            if (!rc) rc = check_hip(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch event");
        if (rc) return rc;
    }
    BatchRing *b = (BatchRing *)ctx->batch_ring;
    const int k = b->next;
    b->next = (k + 1) % BATCH_SLOTS;
    if (b->pending[k]) {
        int rc = check_hip(ctx, hipEventSynchronize(b->ev[k]), "batch slot wait");
        if (rc) return rc;
    }
    const int n = (int)A.size();
    char *h = b->host + BATCH_SLOT_BYTES * k, *d = b->dev + BATCH_SLOT_BYTES * k;
    std::memcpy(h, A.data(), sizeof(StageArgs) * n);
    int *start = (int *)(h + sizeof(StageArgs) * BATCH_MAX);
    int total = 0;
    for (int q = 0; q < n; q++) {
        start[q] = total;
        total += A[q].nbx * ((A[q].nstrip + 3) / 4);
    }
    start[n] = total;
    int rc = check_hip(ctx, hipMemcpyAsync(d, h, BATCH_SLOT_BYTES, hipMemcpyHostToDevice, ctx->stream), "batch upload");
    if (rc) return rc;
    rc = check_hip(ctx, hipEventRecord(b->ev[k], ctx->stream), "batch event record");
    b->pending[k] = rc == VX_OK;
    if (rc) return rc;
    const StageArgs *dA = (const StageArgs *)d;
    const int *dS = (const int *)(d + sizeof(StageArgs) * BATCH_MAX);
    dim3 grid((unsigned)total);
#define VX_BATCH(S)                                                                                   \
    if (mp) k_stage_batch<S, true><<<grid, 256, 0, ctx->stream>>>(dA, dS, n);                        \
    else k_stage_batch<S, false><<<grid, 256, 0, ctx->stream>>>(dA, dS, n);
    switch (step) {
        case 0: VX_BATCH(0) break;
        case 1: VX_BATCH(1) break;
        case 2: VX_BATCH(2) break;
        case 3: VX_BATCH(3) break;
        case 4: VX_BATCH(4) break;
        default: VX_BATCH(5) break;
    }
#undef VX_BATCH
    return check_launch(ctx, "k_stage_batch");
}

extern "C" {

int vx_fct_stage_fused(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl,
                       const double *fin, double *fout) {
    return stage_fused(ctx, p, step, dt, fl, fin, fout, false, 0.0);
}

int vx_fct_stage_fused_mp(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl,
                          const double *fin, double *fout, double f_scale) {
    return stage_fused(ctx, p, step, dt, fl, fin, fout, true, f_scale);
}

int vx_fct_stage_fused_batch(vx_ctx *ctx, const vx_stage_item *items, int n, int step) {
    if (n < 0 || n > BATCH_MAX || (n > 0 && !items))
        return fail(ctx, VX_EINVAL, "stage batch: %d items (at most %d)", n, BATCH_MAX);
    std::vector<StageArgs> fp64, mixed;
    for (int q = 0; q < n; q++) {
        const vx_stage_item &it = items[q];
        StageArgs A;
        const bool mp = it.f_scale > 0.0;
        int rc = stage_args(ctx, &it.p, step, it.dt, &it.fl, it.fin, it.fout, mp, it.f_scale, A);
        if (rc) return rc;
        (mp ? mixed : fp64).push_back(A);
    }
    LaunchTimer lt(ctx, "k_stage");
    int rc = batch_launch(ctx, fp64, step, false);
    return rc ? rc : batch_launch(ctx, mixed, step, true);
}

static int tile_choice(vx_ctx *ctx, int32_t n_x, int32_t n_p, int step, int mixed) {
    for (auto &c : ctx->ti_cache)
        if (c[0] == n_x && c[1] == n_p && c[3] == (mixed ? 1 : 0) && c[4] == step) return c[2];
    return tile_lookup(ctx->device, n_x, n_p, mixed ? 1 : 0, step);
}

int vx_stage_tile_rows(vx_ctx *ctx, int32_t n_x, int32_t n_p, int step, int mixed) {
    return tile_choice(ctx, n_x, n_p, step, mixed) % W4_FORM;
}

int vx_stage_tile_waves(vx_ctx *ctx, int32_t n_x, int32_t n_p, int step, int mixed) {
    const int c = tile_choice(ctx, n_x, n_p, step, mixed);
    return c == 0 ? 0 : (c >= W4_FORM ? 4 : 3);
}

}  // extern "C"
