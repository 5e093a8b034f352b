// Device helpers shared by the FCT, moment and field kernels.  Each follows
// the reference expression for expression (same operands, same association)
// because the FP64 path is bit-compared with the CPU oracle; the library is
// compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>

#include "vx_consts.h"

namespace vxd {

// Rectangle.hpp:111-126
__device__ __forceinline__ double valmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double valmin(double a, double b) { return a < b ? a : b; }
// std::min / std::max as used by the limiter-C sync (Rectangle.cpp:1646-1829)
__device__ __forceinline__ double stdmin(double a, double b) { return (b < a) ? b : a; }

// XCD-aware workgroup order: the dispatcher hands workgroup w of a launch to
// XCD w mod 8, so map w to a tile index that gives each XCD one contiguous
// run of tiles.  Neighbouring tiles (which read each other's halo rows and
// columns) then run on the same XCD and share those cache lines in its L2.
// A bijection of [0, nwg): XCD x takes q + (x < rem) workgroups and the run
// [x q + min(x, rem), ...) of that length.
__device__ __forceinline__ int xcd_order(int w, int nwg) {
    const int xcd = w & 7, q = nwg >> 3, rem = nwg & 7;
    return xcd * q + min(xcd, rem) + (w >> 3);
}

// The same in chunks of K consecutive tiles, dealt round-robin to the XCDs
// (chunk c to XCD c mod 8): neighbours within a chunk share one L2, and
// launches whose tiles differ in cost (several patches of different strip
// shapes) stay balanced over the XCDs.  Workgroups past the last whole round
// of 8 K keep their own index.  A bijection of [0, nwg).
template <int K>
__device__ __forceinline__ int xcd_chunk_order(int w, int nwg) {
    const int full = nwg - nwg % (8 * K);
    if (w >= full) return w;
    const int x = w & 7, k = w >> 3;
    return (x + 8 * (k / K)) * K + k % K;
}

// v_max_f64 / v_min_f64 as single instructions.  fmax/fmin compile to the
// same instruction, but first quiet possible signalling NaNs with a
// v_max_f64 x, x, x of every operand the compiler cannot prove canonical
// (loaded values, lane shifts, values carried around the row loop); the
// operands here are never NaN, so that step changes no bits and only costs
// issue slots.
__device__ __forceinline__ double hwmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double hwmin(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Rectangle.hpp:86-88: lower p-face of cell j (pmin + dp * (j + p_pos))
__device__ __forceinline__ double momentum(double pmin, double dp, int j, int p_pos) {
    return pmin + dp * ((double)j + (double)p_pos);
}

// Rectangle::GetEfield(i) (Rectangle.cpp:1055-1067): mean of the finest-grid E
// over the rtb finest cells of patch column i.  EMFieldSolver::GetEfield
// (EMSolver.cpp:137-154) wraps each stencil index once, which equals the
// periodic index for finest indices in [-n_f, 2 n_f); callers check that the
// patch lies inside the finest grid, so rtb * (i + x_pos) + k stays there for
// the columns i in [-2, n_x + 2) the stage reads.
__device__ __forceinline__ double patch_E(const double *E, int n_f, int rtb, int x_pos, int i) {
    const int j = rtb * (i + x_pos);
    double t = 0.0;
    for (int k = 0; k < rtb; k++) {
        int m = j + k;
        m = m < 0 ? m + n_f : (m >= n_f ? m - n_f : m);
        t += E[m];
    }
    t *= (1.0 / (double)rtb);
    return t;
}

// Rectangle.hpp:181-183 with mc2inv = (m_inv*c_inv)*(m_inv*c_inv)
__device__ __forceinline__ double gamma_(double p, double a2, double mc2inv) {
    return sqrt(1.0 + ((p * p) + a2) * mc2inv);
}

// Rectangle.cpp:979-1030 (two-candidate WENO, upwind side takes the larger
// mapped weight); identical values to GetWenoEdgeValue (Rectangle.cpp:943-977).
__device__ __forceinline__ double weno(double f1, double f2, double f3, double f4, bool right) {
    double fL = (1.0 / 6) * (-f1 + 5 * f2 + 2 * f3);
    double fR = (1.0 / 6) * (2 * f2 + 5 * f3 - f4);
    double AL = f1 - 2 * f2 + f3;
    double BL = f3 - f1;
    double AR = f2 - 2 * f3 + f4;
    double BR = f4 - f2;
    double bL = 4.0 / 3 * (AL * AL) + 0.5 * AL * BL + 0.25 * (BL * BL);
    double bR = 4.0 / 3 * (AR * AR) - 0.5 * AR * BR + 0.25 * (BR * BR);
    const double mm = 1.0e-10;
    double oL = 0.5 / ((mm + bL) * (mm + bL));
    double oR = 0.5 / ((mm + bR) * (mm + bR));
    double wL = oL / (oL + oR);
    double wR = oR / (oL + oR);
    double wL0 = wL * (0.75 + wL * (wL - 0.5));
    double wR0 = wR * (0.75 + wR * (wR - 0.5));
    double s = wL0 + wR0;
    // right: the larger weight goes to fL; left: the smaller one does
    bool takeL = right ? (wL0 > wR0) : (wL0 < wR0);
    double w = (takeL ? wL0 : wR0) / s;
    double fwL = w;
    double fwR = 1 - w;
    return fwL * fL + fwR * fR;
}

// EMSolver.cpp:558-585: plain two-candidate WENO for the face potentials
__device__ __forceinline__ double face_weno(double f1, double f2, double f3, double f4) {
    double fL = (1.0 / 6) * (-f1 + 5 * f2 + 2 * f3);
    double fR = (1.0 / 6) * (2 * f2 + 5 * f3 - f4);
    double AL = f1 - 2 * f2 + f3;
    double BL = f3 - f1;
    double AR = f2 - 2 * f3 + f4;
    double BR = f4 - f2;
    double bL = 4.0 / 3 * (AL * AL) + 0.5 * AL * BL + 0.25 * (BL * BL);
    double bR = 4.0 / 3 * (AR * AR) - 0.5 * AR * BR + 0.25 * (BR * BR);
    const double mm = 1.0e-10;
    double oL = 0.5 / ((mm + bL) * (mm + bL));
    double oR = 0.5 / ((mm + bR) * (mm + bR));
    double wL = oL / (oL + oR);
    double wR = oR / (oL + oR);
    return wL * fL + wR * fR;
}

// Whole-wavefront lane shifts by one lane (DPP wave_shl:1 / wave_shr:1, plain
// VALU moves instead of LDS permutes): shdn gives lane l the value of lane
// l + 1, shup the value of lane l - 1.  The end lane (63, resp. 0) receives 0;
// it is a halo lane, as are its neighbours that consume it.
template <int CTRL>
__device__ __forceinline__ double wave_shift1(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
    const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
    return __hiloint2double(rhi, rlo);
}
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
__device__ __forceinline__ double shdn1(double v) { return wave_shift1<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ double shup1(double v) { return wave_shift1<DPP_WAVE_SHR1>(v); }

// Mixed-precision helpers: f scaled by an exact power of two (sinv = 2^-k
// with 2^k near the patch's max |f|) and rounded to FP32 for the WENO weights
// and the Zalesak ratios; face values, fluxes and updates stay FP64.
// WENO face value with the weights in FP32 (g: f * sinv in FP32) and the
// candidates and the blend in FP64; same candidate choice as vxd::weno.
__device__ __forceinline__ double weno_mp(double f1, double f2, double f3, double f4, float g1, float g2,
                                          float g3, float g4, float eps, bool right) {
    const double fL = (1.0 / 6) * (-f1 + 5 * f2 + 2 * f3);
    const double fR = (1.0 / 6) * (2 * f2 + 5 * f3 - f4);
    const float AL = g1 - 2 * g2 + g3, BL = g3 - g1;
    const float AR = g2 - 2 * g3 + g4, BR = g4 - g2;
    const float bL = eps + (4.0f / 3 * (AL * AL) + 0.5f * AL * BL + 0.25f * (BL * BL));
    const float bR = eps + (4.0f / 3 * (AR * AR) - 0.5f * AR * BR + 0.25f * (BR * BR));
    // oL / (oL + oR) with o = 0.5 / b^2 is bR^2 / (bL^2 + bR^2); both are
    // divided by the larger b first, so the squares stay in [0, 1]
    const float m = fmaxf(bL, bR);
    const bool ok = m > 1e-30f;
    const float rm = __builtin_amdgcn_rcpf(ok ? m : 1.0f);
    const float x = bL * rm, y = bR * rm, x2 = x * x, y2 = y * y;
    const float rden = __builtin_amdgcn_rcpf(x2 + y2);
    const float wL = ok ? y2 * rden : 0.5f, wR = ok ? x2 * rden : 0.5f;
    const float wL0 = wL * (0.75f + wL * (wL - 0.5f));
    const float wR0 = wR * (0.75f + wR * (wR - 0.5f));
    const bool takeL = right ? (wL0 > wR0) : (wL0 < wR0);
    const float w = (takeL ? wL0 : wR0) * __builtin_amdgcn_rcpf(wL0 + wR0);
    return fR + (double)w * (fL - fR);
}

// Zalesak ratio min(1, Q / P) for P > 0 in FP32 on the scaled values
__device__ __forceinline__ double zalesak_mp(double Q, double P, double sinv) {
    const float qf = (float)(Q * sinv), pf = (float)(P * sinv);
    if (pf > 1e-30f) return (double)fminf(1.0f, qf * __builtin_amdgcn_rcpf(pf));
    return qf > 0.0f ? 1.0 : 0.0;
}

// FP32 lane shifts for the scaled copies of f
__device__ __forceinline__ float shdn1f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float shup1f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

}  // namespace vxd
