// CPU ORACLE -- test infrastructure only (see oracle.hpp for the contract).
// Restates Libbum/Veritas's Vlasov-Maxwell time advance in plain FP64 C++.
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off; no FMA contraction so that
// the GPU kernels, compiled the same way, reproduce the bits).
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "oracle_consts.h"
#include "../veritas_amd/csrc/vx_math.h"

namespace orc {

// ---------------------------------------------------------------- Settings
// Settings.cpp:5-58 with the scenario playing initialConditions and
// settingsOverride (veritas.cpp:7-73).
Settings::Settings(const vx_settings_desc &d0, const vx_hooks &h) : hooks(h) {
    if (!h.initial_distribution) throw std::runtime_error("InitialDistribution hook is required");
    vx_settings_desc d = d0;
    // Settings.cpp:31: settingsOverride runs on the copied inputs
    if (h.settings_override) h.settings_override(h.user, &d);
    if (d.n_species < 1 || d.n_species > VX_MAX_SPECIES) throw std::runtime_error("1 to 8 species");
    dx = d.dx;
    dp.assign(d.dp, d.dp + d.n_species);
    maxDepth = (int)d.Lfinest - 1;
    x_size = d.nx;
    p_size.assign(d.np, d.np + d.n_species);
    refinementRatio = d.r;
    refinementCriteria = d.refinementCriteria;
    minEfficiency = d.minEfficiency;
    time = 0.0;
    preLength = d.preLength;
    postLength = d.postLength;
    cfl = d.cfl;
    sizeWeight = d.sizeWeight;
    m.assign(d.mass, d.mass + d.n_species);
    q.assign(d.charge, d.charge + d.n_species);
    pmin.assign(d.pmin, d.pmin + d.n_species);
    quadratureDepth = d.quadratureDepth;
    // Settings.cpp:35-50 hard failures
    if (refinementRatio % 2) throw std::runtime_error("refinementRatio must be even");
    if (x_size % refinementRatio != 0) throw std::runtime_error("x_size not divisible by ratio");
    for (unsigned p : p_size)
        if (p % refinementRatio != 0) throw std::runtime_error("p_size not divisible by ratio");
    // Settings.cpp:53-58
    x_size_finest = x_size * std::pow(refinementRatio, maxDepth);
    for (unsigned i = 0; i < p_size.size(); i++)
        p_size_finest.push_back(p_size[i] * std::pow(refinementRatio, maxDepth));
    fMax.resize(q.size(), 0.0);
    DetermineMaximum();
}

// Settings.cpp:138-152
double Settings::GetDp(int level, int sp) const {
    return std::pow(refinementRatio, level - maxDepth) * dp.at(sp);
}
double Settings::GetDx(int level) const { return std::pow(refinementRatio, level - maxDepth) * dx; }
int Settings::GetXSize(int level) const { return x_size * std::pow(refinementRatio, -level + maxDepth); }
int Settings::GetPSize(int level, int sp) const {
    return p_size[sp] * std::pow(refinementRatio, -level + maxDepth);
}

// Settings.cpp:166-179: the RK stage clock accumulates increments.
void Settings::UpdateTime(int step, double dt) {
    if (step == 1) time += (0.5 * dt);
    else if (step == 2) time += (0.332 - 0.5) * dt;
    else if (step == 3) time += (0.62 - 0.332) * dt;
    else if (step == 4) time += (0.85 - 0.62) * dt;
    else if (step == 5) time += (1.0 - 0.85) * dt;
}

// Settings.cpp:181-195 (mValue deliberately not reset between species)
void Settings::DetermineMaximum() {
    double dxf = GetDx(0);
    double mValue(0.0);
    for (unsigned j = 0; j < q.size(); j++) {
        for (int i = 0; i < GetXSize(0); i++) {
            double t(InitialDistribution((i + 0.5) * dxf, 0, j));
            mValue = t > mValue ? t : mValue;
        }
        fMax[j] = mValue;
    }
}

// ---------------------------------------------------------------- Patch
// Rectangle.cpp:7-113
Patch::Patch(int n_x_, int n_p_, int x_pos_, int p_pos_, int depth_, Settings &s, bool up_,
             bool down_, bool left_, bool right_, int sp_)
    : n_x(n_x_), n_p(n_p_), x_pos(x_pos_), p_pos(p_pos_), depth(depth_), sp(sp_),
      r(s.refinementRatio), up(up_), down(down_), left(left_), right(right_), settings(&s) {
    rtb = std::pow(r, depth);
    dp = s.GetDp(depth, sp);
    dx = s.GetDx(depth);
    m_inv = 1 / s.m.at(sp);
    size_t NSsz = (size_t)(n_x + 4) * (n_p + 4);
    for (auto &v : f) v.assign(NSsz, 0.0);
    for (auto *v : {&ex, &fx, &ep, &fp, &FxLS, &FpLS, &FxDS, &FpDS, &Rp, &Rm, &Cx, &Cp})
        v->assign(NSsz, 0.0);
    for (int k = 0; k < 6; k++) {
        FxL[k].assign(NSsz, 0.0);
        FxH[k].assign(NSsz, 0.0);
        FpL[k].assign(NSsz, 0.0);
        FpH[k].assign(NSsz, 0.0);
    }
    is_nested.assign(NSsz, 0);
    is_interpolated.assign(NSsz, 0);
    is_ilb_x.assign(NSsz, 0);
    is_ilb_p.assign(NSsz, 0);
    finer_level.assign(NSsz, nullptr);
    finer_level_x.assign(NSsz, nullptr);
    finer_level_p.assign(NSsz, nullptr);
    ext_xm.assign(n_p / r + 2, nullptr);
    ext_xp.assign(n_p / r + 2, nullptr);
    ext_pm.assign(n_x / r, nullptr);
    ext_pp.assign(n_x / r, nullptr);
    same_xm.assign(n_p / r + 2, 1);
    same_xp.assign(n_p / r + 2, 1);
    same_pm.assign(n_x / r, 1);
    same_pp.assign(n_x / r, 1);
    chargeR.assign((size_t)(n_x * rtb), 0.0);
    currentR.assign((size_t)(n_x * rtb), 0.0);
    energyR.assign((size_t)(n_p * rtb), 0.0);
    // Rectangle.cpp:68-74 (depth 0 is the finest level)
    errorWeights.assign(5, 0.0);
    double fmx = s.fMax.at(sp);
    errorWeights[0] = 0.5 / fmx * std::pow(r, -0.5 * depth);
    errorWeights[1] = 0.5 / fmx * std::pow(r, -0.5 * depth);
    errorWeights[2] = 0.5 / fmx * std::pow(r, -2.0 * depth);
    errorWeights[3] = 0.5 / fmx * std::pow(r, -2.0 * depth);
    errorWeights[4] = s.sizeWeight / fmx * std::pow(r, -s.maxDepth + depth);
    // Rectangle.cpp:91-101 sub-cell averages for one refinement step
    interpREF.assign(3 * (size_t)r, 0.0);
    for (int i = 0; i < r; i++) {
        double tl = -0.5 + i / (double)r;
        double tr = -0.5 + (i + 1.0) / (double)r;
        interpREF[3 * i] = (tl + tr) * 0.5;
        interpREF[3 * i + 1] = (tl * tl + tl * tr + tr * tr) / 3.0 - (1.0 / 12);
        interpREF[3 * i + 2] = (tl * tl * tl + tl * tl * tr + tl * tr * tr + tr * tr * tr) * 0.25;
    }
    // Rectangle.cpp:103-110 sub-cell average coefficients of the cubic
    interpREL.assign(3 * (size_t)rtb, 0.0);
    for (int i = 0; i < rtb; i++) {
        double tl = -0.5 + i / rtb;
        double tr = -0.5 + (i + 1.0) / rtb;
        interpREL[3 * i] = (tl + tr) * 0.5;
        interpREL[3 * i + 1] = (tl * tl + tl * tr + tr * tr) / 3.0 - (1.0 / 12);
        interpREL[3 * i + 2] = (tl * tl * tl + tl * tl * tr + tl * tr * tr + tr * tr * tr) * 0.25;
    }
}

// Rectangle.hpp:86-88: lower p-face of cell i
double Patch::Momentum(double i) const { return settings->pmin[sp] + dp * (i + p_pos); }

// Rectangle.hpp:181-183
double Patch::Gamma(double p, double a2) const {
    return std::sqrt(1.0 + ((p * p) + a2) * ((m_inv * kCInv) * (m_inv * kCInv)));
}

// Rectangle.cpp:307-312; the sentinel (BoundaryCondition.cpp:6-8) returns 0
double Patch::GetValueFromSameLevel(int i, int j, int val) const {
    return f[val][NS(i - x_pos, j - p_pos)];
}

// Rectangle.cpp:139-155 (REL variant; interpolation matrix in oracle.hpp)

std::vector<double> Patch::GetInterpolantsREL(double f1, double f2, double f3, double f4,
                                              double f5) const {
    int n = (int)rtb;
    std::vector<double> t(n, 0.0);
    f5 -= f3; f4 -= f3; f2 -= f3; f1 -= f3;
    double a1 = kInterp[0] * f1 + kInterp[1] * f2 + kInterp[2] * f4 + kInterp[3] * f5;
    double a2 = kInterp[4] * f1 + kInterp[5] * f2 + kInterp[6] * f4 + kInterp[7] * f5;
    double a3 = kInterp[8] * f1 + kInterp[9] * f2 + kInterp[10] * f4 + kInterp[11] * f5;
    for (int i = 0; i < n; i++)
        t[i] = interpREL[3 * i] * a1 + interpREL[3 * i + 1] * a2 + interpREL[3 * i + 2] * a3 + f3;
    return t;
}

// Rectangle.cpp:616-665 (MKLINIT == 0 branch)
void Patch::InitializeDistribution() {
    const int nvals = std::pow(r, depth + settings->quadratureDepth);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n_x; i++) {
        for (int j = 0; j < n_p; j++) {
            double temp = 0.0, xp = 0.0, pp = 0.0;
            for (int k = 0; k < nvals; k++) {
                for (int l = 0; l < nvals; l++) {
                    xp = ((0.5 + k) / nvals + i + x_pos) * dx;
                    pp = Momentum((0.5 + l) / nvals + j);
                    temp += settings->InitialDistribution(xp, pp, sp);
                }
            }
            f[0][NS(i, j)] = temp * (1 / std::pow(nvals, 2.0));
            f[1][NS(i, j)] = temp * (1 / std::pow(nvals, 2.0));
        }
    }
}

// Rectangle.cpp:1055-1067
double Patch::GetEfield(int i) const {
    int j = GetFinestIndex(i);
    double temp(0.0);
    for (int k = 0; k < rtb; k++) temp += em->GetEfield(j + k);
    temp *= (1.0 / rtb);
    return temp;
}

// Rectangle.cpp:157-282, MKL branch (per-sub-cell gamma), vx_log for vdLn
void Patch::CalculateRhoAndJ() {
    std::fill(chargeR.begin(), chargeR.end(), 0.0);
    std::fill(currentR.begin(), currentR.end(), 0.0);
    const double q = settings->q.at(sp);
    const double mass = settings->m.at(sp);
    const double c1 = m_inv * kCInv;
    const double c2 = 1 / c1;
    const double c3 = 1 / 48.0;
    const int n = (int)rtb;
    const std::vector<double> &g = f[1];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n_x; i++) {
        std::vector<double> lin(3 * n), a2(n);
        int j = 0;
        auto rel = [&](int jj) {
            return GetInterpolantsREL(g[NS(i - 2, jj)], g[NS(i - 1, jj)], g[NS(i, jj)],
                                      g[NS(i + 1, jj)], g[NS(i + 2, jj)]);
        };
        std::vector<double> temp = rel(j), tempm1 = rel(j - 1), tempp1 = rel(j + 1);
        double te = 0.0, tp1 = 0.0, tm1 = 0.0;
        for (int k = 0; k < n; k++) {
            te += temp[k];
            tp1 += tempp1[k];
            tm1 += tempm1[k];
        }
        double cor = g[NS(i, j)] - (1.0 / rtb) * te;
        double corm1 = g[NS(i, j - 1)] - (1.0 / rtb) * tm1;
        double corp1 = g[NS(i, j + 1)] - (1.0 / rtb) * tp1;
        for (int k = 0; k < n; k++) {
            temp[k] += cor;
            tempm1[k] += corm1;
            tempp1[k] += corp1;
        }
        for (j = 0; j < n_p; j++) {
            if (!is_nested[NS(i, j)]) {
                for (int k = 0; k < n; k++) {
                    a2[k] = em->GetCellAverageASquared((i + x_pos) * n + k);
                    chargeR[i * n + k] += temp[k];
                    double gm1 = Gamma(Momentum(j - 1), q * q * a2[k]);
                    double g0 = Gamma(Momentum(j), q * q * a2[k]);
                    double g1 = Gamma(Momentum(j + 1), q * q * a2[k]);
                    double g2 = Gamma(Momentum(j + 2), q * q * a2[k]);
                    lin[k] = (g1 + c1 * Momentum(j + 1)) / (g0 + c1 * Momentum(j));
                    lin[k + n] = (g0 + c1 * Momentum(j)) / (gm1 + c1 * Momentum(j - 1));
                    lin[k + 2 * n] = (g2 + c1 * Momentum(j + 2)) / (g1 + c1 * Momentum(j + 1));
                }
                for (int k = 0; k < n; k++) {
                    double gm = c2 * vx_log(lin[k]);
                    double gmm = c2 * vx_log(lin[k + n]);
                    double gp = c2 * vx_log(lin[k + 2 * n]);
                    currentR[i * n + k] += temp[k] * gm + c3 * (gp - gmm) * (tempp1[k] - tempm1[k]);
                }
            }
            tempm1.swap(temp);
            temp.swap(tempp1);
            tempp1 = rel(j + 2);
            tp1 = 0.0;
            for (int k = 0; k < n; k++) tp1 += tempp1[k];
            corp1 = g[NS(i, j + 2)] - (1.0 / rtb) * tp1;
            for (int k = 0; k < n; k++) tempp1[k] += corp1;
        }
        for (int k = 0; k < n; k++) {
            chargeR[i * n + k] *= dp * q;
            currentR[i * n + k] *= -q * q / mass;
        }
    }
}

// Rectangle.cpp:284-305 without the race: sequential over i then j
void Patch::CalculateEnergy() {
    std::fill(energyR.begin(), energyR.end(), 0.0);
    int n = (int)rtb;
    const std::vector<double> &g = f[1];
    for (int i = 0; i < n_x; i++)
        for (int j = 0; j < n_p; j++)
            if (!is_nested[NS(i, j)]) {
                std::vector<double> t = GetInterpolantsREL(g[NS(i, j - 2)], g[NS(i, j - 1)],
                                                           g[NS(i, j)], g[NS(i, j + 1)],
                                                           g[NS(i, j + 2)]);
                for (int k = 0; k < n; k++) energyR[j * n + k] += t[k];
            }
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n_p; j++) energyR[j * n + k] *= dx;
}

// Rectangle.cpp:979-1030
static double WenoNoMax(double f1, double f2, double f3, double f4, bool right) {
    double fL = (1.0 / 6) * (-f1 + 5 * f2 + 2 * f3);
    double fR = (1.0 / 6) * (2 * f2 + 5 * f3 - f4);
    double AL = f1 - 2 * f2 + f3;
    double BL = f3 - f1;
    double AR = f2 - 2 * f3 + f4;
    double BR = f4 - f2;
    double bL = 4.0 / 3 * (AL * AL) + 0.5 * AL * BL + 0.25 * (BL * BL);
    double bR = 4.0 / 3 * (AR * AR) - 0.5 * AR * BR + 0.25 * (BR * BR);
    double mm(1.0e-10);
    double oL = 0.5 / ((mm + bL) * (mm + bL));
    double oR = 0.5 / ((mm + bR) * (mm + bR));
    double wL = oL / (oL + oR);
    double wR = oR / (oL + oR);
    double wL0 = wL * (0.75 + wL * (wL - 0.5));
    double wR0 = wR * (0.75 + wR * (wR - 0.5));
    double fwL = 0.0, fwR = 0.0;
    if (right) {
        if (wL0 > wR0) {
            fwL = wL0 / (wL0 + wR0);
            fwR = 1 - (wL0 / (wL0 + wR0));
        } else {
            fwL = wR0 / (wL0 + wR0);
            fwR = 1 - (wR0 / (wL0 + wR0));
        }
    } else {
        if (wL0 < wR0) {
            fwL = wL0 / (wL0 + wR0);
            fwR = 1 - (wL0 / (wL0 + wR0));
        } else {
            fwL = wR0 / (wL0 + wR0);
            fwR = 1 - (wR0 / (wL0 + wR0));
        }
    }
    return (fwL * (fL) + fwR * (fR));
}

double weno_nomax(double f1, double f2, double f3, double f4, bool right) {
    return WenoNoMax(f1, f2, f3, f4, right);
}

static inline double valmax(double a, double b) { return a > b ? a : b; }  // Rectangle.hpp:111
static inline double valmin(double a, double b) { return a < b ? a : b; }  // Rectangle.hpp:120

// Rectangle.cpp:1255-1623
void Patch::FCTTimeStep(double timestep, int step, int subStep) {
    int xm = left ? 1 : 0;
    int xp = right ? n_x : (n_x + 1);
    int pp = up ? n_p : (n_p + 1);
    int pm = down ? 1 : 0;
    int nx = n_x, np = n_p, nx1 = n_x + 1, nx2 = n_x + 2, np1 = n_p + 1, np2 = n_p + 2;
    std::vector<double> &f0 = f[0], &f1 = f[1], &f2 = f[2];

    if (subStep == 0) {
        double q = settings->q.at(sp);
        double w3 = 1 / 48.0;
        double dx_inv = 1 / dx;
        double dp_inv = 1 / dp;
        double cc = kC * kC * settings->m.at(sp);
#pragma omp parallel for schedule(static)
        for (int i = -1; i < nx2; i++) {  // 1278-1286
            double as = q * q * em->GetASquared(GetFinestIndex(i));
            for (int j = -1; j < np1; j++)
                ex[NS(i, j)] = dp_inv * cc * (Gamma(Momentum(j + 1), as) - Gamma(Momentum(j), as));
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx1; i++)  // 1287-1293
            for (int j = -1; j < np1; j++)
                fx[NS(i, j)] = WenoNoMax(f1[NS(i - 2, j)], f1[NS(i - 1, j)], f1[NS(i, j)],
                                         f1[NS(i + 1, j)], ex[NS(i, j)] > 0.0);
#pragma omp parallel for schedule(static)
        for (int i = -1; i < nx1; i++) {  // 1294-1305
            double as_1 = q * q * em->GetASquared(GetFinestIndex(i));
            double as_2 = q * q * em->GetASquared(GetFinestIndex(i + 1));
            double Em = q * GetEfield(i);
            for (int j = -1; j < np2; j++) {
                double momentum = Momentum(j);
                ep[NS(i, j)] = Em - cc * dx_inv * (Gamma(momentum, as_2) - Gamma(momentum, as_1));
            }
        }
#pragma omp parallel for schedule(static)
        for (int i = -1; i < nx1; i++)  // 1306-1312
            for (int j = 0; j < np1; j++)
                fp[NS(i, j)] = WenoNoMax(f1[NS(i, j - 2)], f1[NS(i, j - 1)], f1[NS(i, j)],
                                         f1[NS(i, j + 1)], ep[NS(i, j)] > 0.0);
        // 1313-1394; faces on the edge of a finer patch take the finer
        // patch's fluxes (CalculateFluxToCoarse*, Rectangle.hpp:132-178)
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx1; i++)
            for (int j = -1; j < np1; j++) {
                if (!is_ilb_x[NS(i, j)]) {
                    double am = ex[NS(i, j)], ap1 = ex[NS(i, j + 1)], am1 = ex[NS(i, j - 1)];
                    double fm = fx[NS(i, j)], fp1 = fx[NS(i, j + 1)], fm1 = fx[NS(i, j - 1)];
                    FxH[step][NS(i, j)] = dx_inv * (fm * am + w3 * (fp1 - fm1) * (ap1 - am1));
                    FxL[step][NS(i, j)] = dx_inv * ((am > 0.0 ? f1[NS(i - 1, j)] : f1[NS(i, j)]) * am);
                } else {
                    Patch *fine = finer_level_x[NS(i, j)];
                    FxH[step][NS(i, j)] = fine->CalculateFluxToCoarseX(x_pos + i, p_pos + j);
                    FxL[step][NS(i, j)] = fine->CalculateFluxToCoarseXL(x_pos + i, p_pos + j);
                }
            }
#pragma omp parallel for schedule(static)
        for (int i = -1; i < nx1; i++)
            for (int j = 0; j < np1; j++) {
                if (!is_ilb_p[NS(i, j)]) {
                    double am = ep[NS(i, j)], ap1 = ep[NS(i + 1, j)], am1 = ep[NS(i - 1, j)];
                    double fm = fp[NS(i, j)], fp1 = fp[NS(i + 1, j)], fm1 = fp[NS(i - 1, j)];
                    FpH[step][NS(i, j)] = dp_inv * (fm * am + w3 * (fp1 - fm1) * (ap1 - am1));
                    FpL[step][NS(i, j)] = dp_inv * ((am > 0.0 ? f1[NS(i, j - 1)] : f1[NS(i, j)]) * am);
                } else {
                    Patch *fine = finer_level_p[NS(i, j)];
                    FpH[step][NS(i, j)] = fine->CalculateFluxToCoarseP(x_pos + i, p_pos + j);
                    FpL[step][NS(i, j)] = fine->CalculateFluxToCoarsePL(x_pos + i, p_pos + j);
                }
            }
        // 1396-1517: RK combination (a_k = coefficient * dt, aSum summed in order)
        double a[6] = {0, 0, 0, 0, 0, 0};
        int ns = step + 1;
        for (int k = 0; k < ns; k++) a[k] = kRK[step][k] * timestep;
        double aSum = a[0];
        for (int k = 1; k < ns; k++) aSum = aSum + a[k];
#pragma omp parallel for schedule(static)
        for (int i = -2; i < nx2; i++)
            for (int j = -2; j < np2; j++) {
                int idx = NS(i, j);
                FxLS[idx] = aSum * FxL[0][idx];
                FpLS[idx] = aSum * FpL[0][idx];
                double sx = a[0] * FxH[0][idx], sp_ = a[0] * FpH[0][idx];
                for (int k = 1; k < ns; k++) {
                    sx = sx + a[k] * FxH[k][idx];
                    sp_ = sp_ + a[k] * FpH[k][idx];
                }
                FxDS[idx] = sx - FxLS[idx];
                FpDS[idx] = sp_ - FpLS[idx];
            }
        for (int i = 0; i < nx; i++)  // 1518-1524
            for (int j = 0; j < np; j++) f2[NS(i, j)] = f0[NS(i, j)];
        for (int i = xm; i < xp; i++)  // 1526-1534 (serial scatter, this order)
            for (int j = pm; j < pp; j++) {
                f2[NS(i, j)] += FxLS[NS(i, j)];
                f2[NS(i - 1, j)] -= FxLS[NS(i, j)];
                f2[NS(i, j)] += FpLS[NS(i, j)];
                f2[NS(i, j - 1)] -= FpLS[NS(i, j)];
            }
    } else if (subStep == 1) {  // 1536-1594
#pragma omp parallel for schedule(static)
        for (int i = -1; i < nx1; i++) {
            std::vector<double> Ppv(np2), Pmv(np2), Qmv(np2), Qpv(np2);
            for (int j = -1; j < np1; j++) {
                int idx = NS(i, j);
                int i2 = NS(i, j);
                Ppv[j + 1] = valmax(0.0, FxDS[idx]) - valmin(0.0, FxDS[NS(i + 1, j)]) +
                             valmax(0.0, FpDS[idx]) - valmin(0.0, FpDS[NS(i, j + 1)]);
                Pmv[j + 1] = valmax(0.0, FxDS[NS(i + 1, j)]) - valmin(0.0, FxDS[idx]) +
                             valmax(0.0, FpDS[NS(i, j + 1)]) - valmin(0.0, FpDS[idx]);
                double w1ma = valmax(f0[NS(i, j)], f2[i2]);
                double w2ma = valmax(f0[NS(i + 1, j)], f2[NS(i + 1, j)]);
                double w3ma = valmax(f0[NS(i - 1, j)], f2[NS(i - 1, j)]);
                double w4ma = valmax(f0[NS(i, j + 1)], f2[NS(i, j + 1)]);
                double w5ma = valmax(f0[NS(i, j - 1)], f2[NS(i, j - 1)]);
                double wMax = valmax(w1ma, valmax(w2ma, valmax(w3ma, valmax(w4ma, w5ma))));
                double w1mi = valmin(f0[NS(i, j)], f2[i2]);
                double w2mi = valmin(f0[NS(i + 1, j)], f2[NS(i + 1, j)]);
                double w3mi = valmin(f0[NS(i - 1, j)], f2[NS(i - 1, j)]);
                double w4mi = valmin(f0[NS(i, j + 1)], f2[NS(i, j + 1)]);
                double w5mi = valmin(f0[NS(i, j - 1)], f2[NS(i, j - 1)]);
                double wMin = valmin(w1mi, valmin(w2mi, valmin(w3mi, valmin(w4mi, w5mi))));
                Qmv[j + 1] = -wMin + f2[i2];
                Qpv[j + 1] = wMax - f2[i2];
            }
            for (int j = 0; j < np2; j++) {
                int idx = NS(i, j - 1);
                Rp[idx] = Ppv[j] > 0.0 ? valmin(1.0, Qpv[j] / Ppv[j]) : 0.0;
                Rm[idx] = Pmv[j] > 0.0 ? valmin(1.0, Qmv[j] / Pmv[j]) : 0.0;
            }
        }
        std::fill(Cp.begin(), Cp.end(), 1.0);
        std::fill(Cx.begin(), Cx.end(), 1.0);
#pragma omp parallel for schedule(static)
        for (int i = 1; i < nx; i++)
            for (int j = 0; j < np; j++) {
                int idx = NS(i, j), idxi = NS(i - 1, j), idxj = NS(i, j - 1);
                Cx[idx] = FxDS[idx] > 0.0 ? valmin(Rp[idx], Rm[idxi]) : valmin(Rp[idxi], Rm[idx]);
                Cp[idx] = FpDS[idx] > 0.0 ? valmin(Rp[idx], Rm[idxj]) : valmin(Rp[idxj], Rm[idx]);
            }
    } else if (subStep == 2) {  // 1595-1613
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < np; j++) f1[NS(i, j)] = f2[NS(i, j)];
        for (int i = xm; i < xp; i++)
            for (int j = pm; j < pp; j++) {
                int idx = NS(i, j);
                f1[NS(i, j)] += Cx[idx] * FxDS[idx];
                f1[NS(i - 1, j)] -= Cx[idx] * FxDS[idx];
                f1[NS(i, j)] += Cp[idx] * FpDS[idx];
                f1[NS(i, j - 1)] -= Cp[idx] * FpDS[idx];
            }
    } else if (subStep == 3) {  // 1614-1622
        for (int i = -2; i < nx2; i++)
            for (int j = -2; j < np2; j++) f0[NS(i, j)] = f1[NS(i, j)];
    }
}

// Rectangle.cpp:562-614 (same-level ghost copy; sentinel neighbours give 0)
void Patch::UpdateSameLevelBoundaries(int val) {
    std::vector<double> &g = f[val];
    for (unsigned i = 0; i < same_xm.size() - 2; i++)
        if (same_xm[i + 1]) {
            Patch *s = ext_xm[i + 1];
            for (int j = 0; j < r; j++) {
                g[NS(-1, i * r + j)] = s ? s->GetValueFromSameLevel(x_pos - 1, p_pos + i * r + j, val) : 0.0;
                g[NS(-2, i * r + j)] = s ? s->GetValueFromSameLevel(x_pos - 2, p_pos + i * r + j, val) : 0.0;
            }
        }
    for (unsigned i = 0; i < same_xp.size() - 2; i++)
        if (same_xp[i + 1]) {
            Patch *s = ext_xp[i + 1];
            for (int j = 0; j < r; j++) {
                g[NS(n_x, i * r + j)] = s ? s->GetValueFromSameLevel(x_pos + n_x, p_pos + i * r + j, val) : 0.0;
                g[NS(n_x + 1, i * r + j)] = s ? s->GetValueFromSameLevel(x_pos + n_x + 1, p_pos + i * r + j, val) : 0.0;
            }
        }
    for (unsigned i = 0; i < same_pm.size(); i++)
        if (same_pm[i]) {
            Patch *s = ext_pm[i];
            for (int j = 0; j < r; j++) {
                g[NS(i * r + j, -1)] = s ? s->GetValueFromSameLevel(x_pos + i * r + j, p_pos - 1, val) : 0.0;
                g[NS(i * r + j, -2)] = s ? s->GetValueFromSameLevel(x_pos + i * r + j, p_pos - 2, val) : 0.0;
            }
        }
    for (unsigned i = 0; i < same_pp.size(); i++)
        if (same_pp[i]) {
            Patch *s = ext_pp[i];
            for (int j = 0; j < r; j++) {
                g[NS(i * r + j, n_p)] = s ? s->GetValueFromSameLevel(x_pos + i * r + j, p_pos + n_p, val) : 0.0;
                g[NS(i * r + j, n_p + 1)] = s ? s->GetValueFromSameLevel(x_pos + i * r + j, p_pos + n_p + 1, val) : 0.0;
            }
        }
}

// ---------------------------------------------------------------- Field
// EMSolver.cpp:6-89
Field::Field(Settings &s, std::vector<Mesh *> ms)
    : settings(s), x_size(s.x_size_finest), meshes(ms), Ex0(0.0) {
    n_prepad = std::max(std::floor(s.preLength / s.GetDx(0)), 2.0);
    n_postpad = std::max(std::floor(s.postLength / s.GetDx(0)), 2.0);
    charge.assign(x_size, 0.0);
    J.assign(x_size, 0.0);
    neutralizationCharge.assign(x_size, 0.0);
    charges.assign(s.q.size(), charge);
    for (unsigned i = 0; i < s.p_size.size(); i++) energies.push_back(std::vector<double>(s.p_size_finest[i], 0.0));
    size_t ne = (size_t)(x_size + n_prepad + n_postpad) * 8;
    for (auto *v : {&By, &Bz, &Ey, &Ez, &Ay, &Az}) v->assign(ne, 0.0);
    a_squared.assign(x_size + 1, 0.0);
    PHI.assign(x_size, 0.0);
    // EMSolver.cpp:28 stencil weights
    w1 = -16.0 / 30.0 / 0.4;
    w2 = 1.0 / 30.0 / 0.4;
    w3 = 1 / 0.4;
    fieldCoef = (1.0 / (12 * s.GetDx(0)));

    // Bordered band LU of A' = rows/cols 1..N-1 of the matrix of EMSolver.cpp:41-67
    // (periodic 5-point stencil with column 0 removed).  B = leading (N-2)x(N-2)
    // pentadiagonal block, border = last unknown (couples to 1 through the wrap).
    int N = (int)x_size, M = N - 2;
    if (N < 8) throw std::runtime_error("Poisson: x_size must be >= 8");
    auto Bent = [&](int r_, int c_) -> double {
        int d = std::abs(r_ - c_);
        return d == 0 ? w3 : d == 1 ? w1 : d == 2 ? w2 : 0.0;
    };
    bandL1.assign(M, 0.0); bandL2.assign(M, 0.0); bandD.assign(M, 0.0);
    bandU1.assign(M, 0.0); bandU2.assign(M, 0.0);
    for (int k = 0; k < M; k++) {
        double d = Bent(k, k);
        if (k >= 1) d -= bandL1[k] * bandU1[k - 1];
        if (k >= 2) d -= bandL2[k] * bandU2[k - 2];
        bandD[k] = d;
        double u1 = (k + 1 < M) ? Bent(k, k + 1) : 0.0;
        if (k >= 1 && k + 1 < M) u1 -= bandL1[k] * bandU2[k - 1];
        bandU1[k] = u1;
        bandU2[k] = (k + 2 < M) ? Bent(k, k + 2) : 0.0;
        if (k + 1 < M) {
            double l = Bent(k + 1, k);
            if (k >= 1) l -= bandL2[k + 1] * bandU1[k - 1];
            bandL1[k + 1] = l / d;
        }
        if (k + 2 < M) bandL2[k + 2] = Bent(k + 2, k) / d;
    }
    // Note: bandL2[k+2] must be known before bandL1[k+1] uses it at k+1; the
    // loop above sets bandL2[k+1] at iteration k-1, so the order is valid.
    std::vector<double> u(M, 0.0), v(M, 0.0);
    u[0] = w2; u[M - 2] += w2; u[M - 1] += w1;  // column N-1, rows 1, N-3, N-2
    v[0] = w2; v[M - 2] += w2; v[M - 1] += w1;  // row N-1, columns 1, N-3, N-2
    Binv_u = u;
    auto bsolve = [&](std::vector<double> &y) {
        for (int k = 0; k < M; k++) {
            if (k >= 1) y[k] -= bandL1[k] * y[k - 1];
            if (k >= 2) y[k] -= bandL2[k] * y[k - 2];
        }
        for (int k = M - 1; k >= 0; k--) {
            if (k + 1 < M) y[k] -= bandU1[k] * y[k + 1];
            if (k + 2 < M) y[k] -= bandU2[k] * y[k + 2];
            y[k] /= bandD[k];
        }
    };
    bsolve(Binv_u);
    double sdot = 0.0;
    for (int k = 0; k < M; k++) sdot += v[k] * Binv_u[k];
    schur = w3 - sdot;
}

// EMSolver.cpp:104-122
void Field::AssembleRhoAndJ() {
    std::fill(charge.begin(), charge.end(), 0.0);
    std::fill(J.begin(), J.end(), 0.0);
    for (unsigned i = 0; i < meshes.size(); i++) std::fill(charges[i].begin(), charges[i].end(), 0.0);
    for (unsigned i = 0; i < meshes.size(); i++) meshes[i]->InterpolateRhoAndJToFinestMesh(charges[i], J);
    for (unsigned i = 0; i < meshes.size(); i++)
        for (unsigned j = 0; j < x_size; j++) charge[j] += charges[i][j];
}

// EMSolver.cpp:124-131
void Field::AssembleEnergy() {
    for (unsigned i = 0; i < meshes.size(); i++) std::fill(energies[i].begin(), energies[i].end(), 0.0);
    for (unsigned i = 0; i < meshes.size(); i++) meshes[i]->InterpolateEnergyToFinestMesh(energies[i]);
}

double Field::GetASquared(int i) const { return a_squared[std::min(std::max(i, 0), (int)x_size)]; }

// EMSolver.cpp:137-154
double Field::GetEfield(int i) const {
    int N = (int)x_size;
    int ip1 = i + 1, im1 = i - 1, ip2 = i + 2, im2 = i - 2;
    ip1 = (ip1 > -1) ? ip1 : (ip1 + N);
    im1 = (im1 > -1) ? im1 : (im1 + N);
    ip2 = (ip2 > -1) ? ip2 : (ip2 + N);
    im2 = (im2 > -1) ? im2 : (im2 + N);
    ip1 = (ip1 < N) ? ip1 : (ip1 - N);
    im1 = (im1 < N) ? im1 : (im1 - N);
    ip2 = (ip2 < N) ? ip2 : (ip2 - N);
    im2 = (im2 < N) ? im2 : (im2 - N);
    return -fieldCoef * (8 * (PHI[ip1] - PHI[im1]) - PHI[ip2] + PHI[im2]) + Ex0;
}

// EMSolver.hpp:57-64
double Field::GetCellAverageASquared(int i) const {
    int Ne = (int)x_size + n_prepad + n_postpad;
    i += n_prepad;
    i = i > -1 ? i : 0;
    i = i < Ne ? i : (Ne - 1);
    double ay = Ay[Index(i, 1)], az = Az[Index(i, 1)];
    return (ay * ay) + (az * az);
}

// Approximate solve of A x = b (A as built at EMSolver.cpp:41-67):
// x_1..x_{N-1} from the bordered band factorisation, x_0 from row 0.
void Field::PoissonApprox(const std::vector<double> &b, std::vector<double> &x) const {
    int N = (int)x_size, M = N - 2;
    std::vector<double> z(b.begin() + 1, b.begin() + 1 + M);
    for (int k = 0; k < M; k++) {
        if (k >= 1) z[k] -= bandL1[k] * z[k - 1];
        if (k >= 2) z[k] -= bandL2[k] * z[k - 2];
    }
    for (int k = M - 1; k >= 0; k--) {
        if (k + 1 < M) z[k] -= bandU1[k] * z[k + 1];
        if (k + 2 < M) z[k] -= bandU2[k] * z[k + 2];
        z[k] /= bandD[k];
    }
    // border row: v = (w2 at 0, w2 at M-2, w1 at M-1)
    double vz = w2 * z[0] + w2 * z[M - 2] + w1 * z[M - 1];
    double xl = (b[N - 1] - vz) / schur;
    x.assign(N, 0.0);
    for (int k = 0; k < M; k++) x[k + 1] = z[k] - Binv_u[k] * xl;
    x[N - 1] = xl;
    x[0] = b[0] - (((w1 * x[1] + w2 * x[2]) + w2 * x[N - 2]) + w1 * x[N - 1]);
}

// Residual b - A x evaluated in double-double (exact products by fma,
// error-free sums), rounded once.  Term order: row 0: x0, x1, x2, x_{N-2},
// x_{N-1}; row i >= 1: x_{i-2}, x_{i-1}, x_i, x_{i+1}, x_{i+2} (mod N), with
// column 0 absent.
void Field::PoissonResidual(const std::vector<double> &b, const std::vector<double> &x,
                            std::vector<double> &res) const {
    int N = (int)x_size;
    res.assign(N, 0.0);
    for (int i = 0; i < N; i++) {
        double sh = b[i], sl = 0.0;
        auto sub = [&](double a, double xv) {
            double p = a * xv;
            double e = std::fma(a, xv, -p);
            double t = sh - p;
            double bv = t - sh;
            double err = (sh - (t - bv)) + (-p - bv);
            sl = sl + (err - e);
            sh = t;
        };
        if (i == 0) {
            sub(1.0, x[0]); sub(w1, x[1]); sub(w2, x[2]); sub(w2, x[N - 2]); sub(w1, x[N - 1]);
        } else {
            const double c[5] = {w2, w1, w3, w1, w2};
            for (int d = -2; d <= 2; d++) {
                int j = i + d;
                j = j < 0 ? j + N : (j >= N ? j - N : j);
                if (j == 0) continue;
                sub(c[d + 2], x[j]);
            }
        }
        res[i] = sh + sl;
    }
}

void Field::PoissonSolve(const std::vector<double> &b, std::vector<double> &x) const {
    PoissonApprox(b, x);
    std::vector<double> res, c;
    for (int it = 0; it < 3; it++) {
        PoissonResidual(b, x, res);
        PoissonApprox(res, c);
        for (unsigned i = 0; i < x_size; i++) x[i] = x[i] + c[i];
    }
}

// EMSolver.cpp:156-192
void Field::UpdatePotential() {
    double te = kEps0Inv;
    std::vector<double> b(x_size);
    for (unsigned i = 0; i < x_size; i++) b[i] = te * (charge[i] + neutralizationCharge[i]);
    double w = std::pow(settings.GetDx(0), 2.0);
    for (unsigned i = 0; i < x_size; i++) b[i] *= w;
    PoissonSolve(b, PHI);
    Ex0 += -(GetEfield(-1) + GetEfield(0)) * 0.5;
}

// EMSolver.cpp:194-202
void Field::RGKStep(int step, double timestep) {
    RGKCalculateRHS(step);
    RGKUpdateIntermediateSolution(step, timestep);
    InterpolateToFaces();
}

// EMSolver.cpp:204-338
void Field::RGKUpdateIntermediateSolution(int step, double timestep) {
    unsigned iloop = x_size + n_prepad + n_postpad;
    std::vector<double> *F[6] = {&By, &Bz, &Ey, &Ez, &Ay, &Az};
    if (step >= 0 && step <= 4) {
        double a[5];
        for (int k = 0; k <= step; k++) a[k] = kRK[step][k] * timestep;
        for (unsigned i = 0; i < iloop; i++)
            for (auto *v : F) {
                std::vector<double> &y = *v;
                double t = y[Index(i, 0)] + a[0] * y[Index(i, 2)];
                for (int k = 1; k <= step; k++) t = t + a[k] * y[Index(i, k + 2)];
                y[Index(i, 1)] = t;
            }
    } else if (step == 5) {
        double a[6], b[6];
        for (int k = 0; k < 5; k++) a[k] = kRK[4][k] * timestep;
        a[5] = 0;
        for (int k = 0; k < 6; k++) b[k] = kRK[5][k] * timestep - a[k];
        for (unsigned i = 0; i < iloop; i++)
            for (auto *v : F) {
                std::vector<double> &y = *v;
                double t = y[Index(i, 1)] + b[0] * y[Index(i, 2)];
                for (int k = 1; k < 6; k++) t = t + b[k] * y[Index(i, k + 2)];
                y[Index(i, 0)] = t;
                y[Index(i, 1)] = y[Index(i, 0)];
            }
    }
}

// EMSolver.cpp:479-553
void Field::RGKCalculateRHS(int step) {
    int iloop = (int)x_size + n_prepad + n_postpad;
    for (int i = 0; i < iloop; i++) {
        int s = Index(i, step + 2);
        By[s] = 0; Bz[s] = 0; Ey[s] = 0; Ez[s] = 0; Ay[s] = 0; Az[s] = 0;
    }
    double dx_inv = 1 / settings.GetDx(0);
    {
        int i = 0;
        int s = Index(i, step + 2), i1 = Index(i, 1), ip1 = Index(i + 1, 1);
        By[s] = dx_inv * (Ez[ip1] - Ez[i1]);
        Bz[s] = -dx_inv * (Ey[ip1] - Ey[i1]);
        Ey[s] = -dx_inv * (Bz[i1] - settings.GetBZ(0, settings.time)) * kEps0Inv * kMuInv;
        Ez[s] = dx_inv * (By[i1] - settings.GetBY(0, settings.time)) * kEps0Inv * kMuInv;
        Ay[s] = -Ey[i1];
        Az[s] = -Ez[i1];
    }
    for (int i = 1; i < n_prepad; i++) {
        int s = Index(i, step + 2), i1 = Index(i, 1), ip1 = Index(i + 1, 1), in1 = Index(i - 1, 1);
        By[s] = dx_inv * (Ez[ip1] - Ez[i1]);
        Bz[s] = -dx_inv * (Ey[ip1] - Ey[i1]);
        Ey[s] = -dx_inv * (Bz[i1] - Bz[in1]) * kEps0Inv * kMuInv;
        Ez[s] = dx_inv * (By[i1] - By[in1]) * kEps0Inv * kMuInv;
        Ay[s] = -Ey[i1];
        Az[s] = -Ez[i1];
    }
    double c1 = -1.0 / 24;
    double c2 = 9.0 / 8.0;
    for (int i = n_prepad; i < n_prepad + (int)x_size; i++) {
        int s = Index(i, step + 2), i1 = Index(i, 1), ip1 = Index(i + 1, 1), in1 = Index(i - 1, 1);
        int ip2 = Index(i + 2, 1), in2 = Index(i - 2, 1);
        By[s] = dx_inv * (c1 * (Ez[ip2] - Ez[in1]) + c2 * (Ez[ip1] - Ez[i1]));
        Bz[s] = -dx_inv * (c1 * (Ey[ip2] - Ey[in1]) + c2 * (Ey[ip1] - Ey[i1]));
        Ey[s] = -dx_inv * (c1 * (Bz[ip1] - Bz[in2]) + c2 * (Bz[i1] - Bz[in1])) * kEps0Inv * kMuInv -
                kEps0Inv * J[i - n_prepad] * Ay[i1];
        Ez[s] = dx_inv * (c1 * (By[ip1] - By[in2]) + c2 * (By[i1] - By[in1])) * kEps0Inv * kMuInv -
                kEps0Inv * J[i - n_prepad] * Az[i1];
        Ay[s] = -Ey[i1];
        Az[s] = -Ez[i1];
    }
    for (int i = n_prepad + (int)x_size; i < iloop - 2; i++) {
        int s = Index(i, step + 2), i1 = Index(i, 1), ip1 = Index(i + 1, 1), in1 = Index(i - 1, 1);
        By[s] = dx_inv * (Ez[ip1] - Ez[i1]);
        Bz[s] = -dx_inv * (Ey[ip1] - Ey[i1]);
        Ey[s] = -dx_inv * (Bz[i1] - Bz[in1]) * kEps0Inv * kMuInv;
        Ez[s] = dx_inv * (By[i1] - By[in1]) * kEps0Inv * kMuInv;
        Ay[s] = -Ey[i1];
        Az[s] = -Ez[i1];
    }
}

// EMSolver.cpp:555-619 (a_squared[N] is never written and stays 0)
static double face_weno_plain(double f1, double f2, double f3, double f4) {
    double fL = (1.0 / 6) * (-f1 + 5 * f2 + 2 * f3);
    double fR = (1.0 / 6) * (2 * f2 + 5 * f3 - f4);
    double AL = f1 - 2 * f2 + f3;
    double BL = f3 - f1;
    double AR = f2 - 2 * f3 + f4;
    double BR = f4 - f2;
    double bL = 4.0 / 3 * (AL * AL) + 0.5 * AL * BL + 0.25 * (BL * BL);
    double bR = 4.0 / 3 * (AR * AR) - 0.5 * AR * BR + 0.25 * (BR * BR);
    double mm(1.0e-10);
    double oL = 0.5 / ((mm + bL) * (mm + bL));
    double oR = 0.5 / ((mm + bR) * (mm + bR));
    double wL = oL / (oL + oR);
    double wR = oR / (oL + oR);
    return wL * fL + wR * fR;
}

double face_weno(double f1, double f2, double f3, double f4) {
    return face_weno_plain(f1, f2, f3, f4);
}

void Field::InterpolateToFaces() {
    for (unsigned i = 0; i < x_size; i++) {
        double a_y = face_weno_plain(Ay[Index(n_prepad + i - 2, 1)], Ay[Index(n_prepad + i - 1, 1)],
                                     Ay[Index(n_prepad + i, 1)], Ay[Index(n_prepad + i + 1, 1)]);
        double a_z = face_weno_plain(Az[Index(n_prepad + i - 2, 1)], Az[Index(n_prepad + i - 1, 1)],
                                     Az[Index(n_prepad + i, 1)], Az[Index(n_prepad + i + 1, 1)]);
        a_squared[i] = (a_y * a_y) + (a_z * a_z);
    }
}

// EMSolver.cpp:621-629
void Field::EnforceChargeNeutralization() {
    AssembleRhoAndJ();
    std::fill(J.begin(), J.end(), 0.0);
    for (unsigned i = 0; i < charge.size(); i++) neutralizationCharge[i] = -charge[i];
}

// EMSolver.cpp:631-664 (field slot 0 indexed without the n_prepad offset, as
// in the reference)
double Field::EstimateCFLBound() const {
    int n = (int)meshes.size();
    const std::vector<double> &m = settings.m, &q = settings.q;
    std::vector<double> dps(n, 0.0);
    double dpsMax(0.0);
    for (int i = 0; i < n; i++) {
        dps[i] = 1 / settings.GetDp(0, i);
        dpsMax = std::max(dpsMax, std::fabs(q[i]) * dps[i]);
    }
    double pc(0.0);
    for (unsigned i = 0; i < x_size; i++) {
        double Azv = Az[i], Ayv = Ay[i];
        double As = Ayv * Ayv + Azv * Azv;
        double temp(0.0);
        for (int j = 0; j < n; j++)
            temp = std::max(temp, std::fabs(q[j]) / m[j] / std::sqrt(1 + As / ((m[j] * kC) * (m[j] * kC))) * dps[j]);
        pc = std::max(pc, temp * std::fabs(Ayv * Bz[i] - Azv * By[i]) + dpsMax * std::fabs(GetEfield(i)));
    }
    return 1.0 / std::max((kC / settings.GetDx(0) + pc), 1e-40);
}

// ---------------------------------------------------------------- Level
Level::Level(int sp_, int depth_, Settings &s) : depth(depth_), sp(sp_), settings(s) {
    chargeL.assign(s.x_size_finest, 0.0);
    energyL.assign(s.p_size_finest[sp], 0.0);
    currentL.assign(s.x_size_finest, 0.0);
}

void Level::FCTTimeStep(double dt, int step, int sub) {  // Level.cpp:12-17
    for (auto &r : rectangles) r->FCTTimeStep(dt, step, sub);
}

void Level::CollectRhoAndJ() {  // Level.cpp:42-62
    for (auto &r : rectangles) r->CalculateRhoAndJ();
    std::fill(chargeL.begin(), chargeL.end(), 0.0);
    std::fill(currentL.begin(), currentL.end(), 0.0);
    for (auto &r : rectangles) {
        int shift = r->x_pos;
        int rtb = (int)r->rtb;
        for (unsigned j = 0; j < r->chargeR.size(); j++) {
            chargeL[shift * rtb + j] += r->chargeR[j];
            currentL[shift * rtb + j] += r->currentR[j];
        }
    }
}

void Level::InterpolateRhoAndJToFinestMesh(std::vector<double> &charge, std::vector<double> &J) {
    CollectRhoAndJ();  // Level.cpp:19-29
    for (unsigned i = 0; i < settings.x_size_finest; i++) {
        charge[i] += chargeL[i];
        J[i] += currentL[i];
    }
}

void Level::CollectEnergy() {  // Level.cpp:64-78
    for (auto &r : rectangles) r->CalculateEnergy();
    std::fill(energyL.begin(), energyL.end(), 0.0);
    for (auto &r : rectangles) {
        int shift = r->p_pos, rtb = (int)r->rtb;
        for (unsigned i = 0; i < r->energyR.size(); i++) energyL[shift * rtb + i] += r->energyR[i];
    }
}

void Level::InterpolateEnergyToFinestMesh(std::vector<double> &energy) {
    CollectEnergy();
    for (unsigned i = 0; i < settings.p_size_finest[sp]; i++) energy[i] += energyL[i];
}

void Level::PushData(int type, int val) {  // Level.cpp:88-126
    switch (type) {
        case 0: for (auto &r : rectangles) r->UpdateInterriorPoints(val); break;
        case 1: for (auto &r : rectangles) r->UpdateSameLevelBoundaries(val); break;
        case 2: for (auto &r : rectangles) r->UpdateDifferentLevelBoundaries(val); break;
        case 3: for (auto &r : rectangles) r->UpdateCornerPoints(val); break;
        case 4: for (auto &r : rectangles) r->CalculateSameBoundaryC(); break;
        case 5: for (auto &r : rectangles) r->CalculateDifferentBoundaryC(); break;
        case 6: for (auto &r : rectangles) r->UpdateSameBoundaryC(); break;
    }
}

// ---------------------------------------------------------------- Mesh
void Mesh::SetFieldSolver(Field *fld) {  // Mesh.cpp:108-114
    em = fld;
    for (auto &lvl : levels)
        for (auto &r : lvl->rectangles) r->em = fld;
}

void Mesh::InterpolateRhoAndJToFinestMesh(std::vector<double> &charge, std::vector<double> &J) {
    for (auto &lvl : levels) lvl->InterpolateRhoAndJToFinestMesh(charge, J);
}

void Mesh::InterpolateEnergyToFinestMesh(std::vector<double> &energy) {
    for (auto &lvl : levels) lvl->InterpolateEnergyToFinestMesh(energy);
}

// Mesh.cpp:64-89
void Mesh::Advance(double dt, int step) {
    for (int i = (int)levels.size() - 1; i > -1; i--) levels[i]->FCTTimeStep(dt, step, 0);
    PushData(2);
    for (int i = (int)levels.size() - 1; i > -1; i--) levels[i]->FCTTimeStep(dt, step, 1);
    PushBoundaryC();
    for (int i = (int)levels.size() - 1; i > -1; i--) levels[i]->FCTTimeStep(dt, step, 2);
    PushData(1);
    if (step == 5)
        for (int i = (int)levels.size() - 1; i > -1; i--) levels[i]->FCTTimeStep(dt, step, 3);
}

// Mesh.cpp:91-106
void Mesh::PushData(int val) {
    levels[0]->PushData(1, val);
    for (unsigned i = 1; i < levels.size(); i++) {
        levels[i]->PushData(0, val);
        levels[i]->PushData(1, val);
    }
    levels[levels.size() - 1]->PushData(3, val);
    for (unsigned i = levels.size() - 1; i > 0; i--) levels[i - 1]->PushData(2, val);
}

// Mesh.cpp:904-917
void Mesh::PushBoundaryC() {
    for (auto &l : levels) l->PushData(4, 1);
    for (auto &l : levels) l->PushData(5, 1);
    for (auto &l : levels) l->PushData(6, 1);
}

// ---------------------------------------------------------------- Solver
// SolverManager.cpp:8-26
Solver::Solver(const vx_settings_desc &d, const vx_hooks &h) : settings(d, h) {
    int nsp = (int)settings.p_size.size();
    for (int i = 0; i < nsp; i++) meshes.push_back(std::unique_ptr<Mesh>(new Mesh(i, settings)));
    std::vector<Mesh *> mp;
    for (auto &m : meshes) mp.push_back(m.get());
    em.reset(new Field(settings, mp));
    for (auto &m : meshes) {
        m->SetFieldSolver(em.get());
        m->PushData();
    }
    em->EnforceChargeNeutralization();
}

// SolverManager.cpp:28-39
void Solver::Advance(double dt) {
    for (int i = 0; i < 6; i++) {
        em->AssembleRhoAndJ();
        em->UpdatePotential();
        for (auto &m : meshes) m->Advance(dt, i);
        settings.UpdateTime(i, dt);
        em->RGKStep(i, dt);
    }
}

// SolverManager.cpp:41-46
void Solver::AdvanceFields(double dt) {
    for (int i = 0; i < 6; i++) {
        settings.UpdateTime(i, dt);
        em->RGKStep(i, dt);
    }
}

}  // namespace orc
