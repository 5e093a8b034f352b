// CPU ORACLE -- test infrastructure only (see oracle.hpp).
//
// Adaptive-mesh part of the restatement: hierarchy construction and regrid
// (Mesh.cpp), patch connectivity, coarse-fine ghost fill, restriction,
// coarse-fine fluxes, limiter-C synchronisation and regrid data transfer
// (Rectangle.cpp, Rectangle.hpp, Level.cpp).  Loop ranges, operation order and
// association follow the reference literally; the domain-edge sentinel
// (BoundaryCondition) is a null neighbour whose same-level value is 0.
//
// Pinned by tests/: restriction = mean of the finer cells, coarse faces on a
// finer edge carry the mean of the finer fluxes, conservation across levels
// (test_oracle_amr.py); clustering by hand-derived boxes (test_clustering.py);
// the 2-D coarse-to-fine ghost fill, corner blocks, same-level copies and the
// regrid transfer by exact fine-cell averages of bilinear data and by
// per-coarse-cell conservation (test_oracle_amr_fill.py).  Parity unpinned:
// the neighbour tables of arrangements those cases do not reach, and the
// limiter-C synchronisation beyond its one-C-per-face invariant.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iterator>
#include <limits>
#include <stdexcept>

#include "oracle_consts.h"
#include "oracle.hpp"

namespace orc {

// ---------------------------------------------------------------- Patch

// Rectangle.cpp:121-137
std::vector<double> Patch::GetInterpolantsREF(double f1, double f2, double f3, double f4,
                                              double f5) const {
    std::vector<double> t(r, 0.0);
    f5 -= f3; f4 -= f3; f2 -= f3; f1 -= f3;
    double a1 = kInterp[0] * f1 + kInterp[1] * f2 + kInterp[2] * f4 + kInterp[3] * f5;
    double a2 = kInterp[4] * f1 + kInterp[5] * f2 + kInterp[6] * f4 + kInterp[7] * f5;
    double a3 = kInterp[8] * f1 + kInterp[9] * f2 + kInterp[10] * f4 + kInterp[11] * f5;
    for (int i = 0; i < r; i++)
        t[i] = interpREF[3 * i] * a1 + interpREF[3 * i + 1] * a2 + interpREF[3 * i + 2] * a3 + f3;
    return t;
}

// Rectangle.cpp:314-327: mean of the r x r fine cells under coarse cell (i, j)
double Patch::GetValueFromFinerLevel(int i, int j, int val) const {
    int i_finer = i * r - x_pos;
    int j_finer = j * r - p_pos;
    double temp(0.0);
    for (int k = 0; k < r; k++)
        for (int l = 0; l < r; l++) temp += f[val][NS(i_finer + k, j_finer + l)];
    temp /= std::pow((double)r, 2.0);
    return temp;
}

// Rectangle.cpp:329-337
void Patch::UpdateInterriorPoints(int val) {
    for (int i = 0; i < n_x; i++)
        for (int j = 0; j < n_p; j++)
            if (is_nested[NS(i, j)])
                f[val][NS(i, j)] = finer_level[NS(i, j)]->GetValueFromFinerLevel(x_pos + i, p_pos + j, val);
}

// Rectangle.cpp:343-415: conservative tensor-product interpolation of the
// coarse cell containing fine cell (i, j); d selects the two sub-cell layers
// next to a fine patch edge (0: below, 1: right, 2: above, 3: left).
std::vector<double> Patch::GetWenoValueFromCoarseLevel(int i, int j, int d, int val) const {
    int i_coarse = i / r - x_pos;
    int j_coarse = j / r - p_pos;
    const std::vector<double> &g = f[val];
    std::vector<std::vector<double>> temps;
    for (int k = -2; k < 3; k++)
        temps.push_back(GetInterpolantsREF(g[NS(i_coarse - 2, j_coarse + k)], g[NS(i_coarse - 1, j_coarse + k)],
                                           g[NS(i_coarse, j_coarse + k)], g[NS(i_coarse + 1, j_coarse + k)],
                                           g[NS(i_coarse + 2, j_coarse + k)]));
    std::vector<double> interpolants(r * r);
    double sum(0.0);
    for (int k = 0; k < r; k++) {
        std::vector<double> part =
            GetInterpolantsREF(temps[0][k], temps[1][k], temps[2][k], temps[3][k], temps[4][k]);
        for (int l = 0; l < r; l++) {
            interpolants[k * r + l] = part[l];
            sum += part[l];
        }
    }
    double correction = g[NS(i_coarse, j_coarse)] - 1.0 / std::pow(r, 2) * sum;
    for (int k = 0; k < r; k++)
        for (int l = 0; l < r; l++) interpolants[k * r + l] += correction;
    std::vector<double> red(2 * r);
    if (d == 0) {
        for (int k = 0; k < r; k++) {
            red[2 * k] = interpolants[r * k + r - 1];
            red[2 * k + 1] = interpolants[r * k + r - 2];
        }
        return red;
    } else if (d == 1) {
        for (int k = 0; k < r; k++) {
            red[2 * k] = interpolants[k];
            red[2 * k + 1] = interpolants[r + k];
        }
        return red;
    } else if (d == 2) {
        for (int k = 0; k < r; k++) {
            red[2 * k] = interpolants[r * k];
            red[2 * k + 1] = interpolants[r * k + 1];
        }
        return red;
    } else if (d == 3) {
        for (int k = 0; k < r; k++) {
            red[2 * k] = interpolants[r * (r - 1) + k];
            red[2 * k + 1] = interpolants[r * (r - 2) + k];
        }
        return red;
    }
    return interpolants;
}

namespace {
// The four x-edge corner blocks shared by UpdateDifferentLevelBoundaries
// (Rectangle.cpp:476-559) and UpdateCornerPoints (Rectangle.cpp:1130-1214).
void corner_blocks(Patch &P, int val) {
    std::vector<double> &g = P.f[val];
    const int r = P.r;
    auto same = [&](Patch *s, int ii, int jj) { return s ? s->GetValueFromSameLevel(ii, jj, val) : 0.0; };
    {
        int i = -1;
        Patch *s = P.ext_xm[i + 1];
        if (!P.same_xm[i + 1]) {
            std::vector<double> t = s->GetWenoValueFromCoarseLevel(P.x_pos - 1, P.p_pos + i * r, 3, val);
            for (int j = r - 2; j < r; j++) {
                g[P.NS(-1, i * r + j)] = t[2 * j];
                g[P.NS(-2, i * r + j)] = t[2 * j + 1];
            }
        } else {
            for (int j = r - 2; j < r; j++) {
                g[P.NS(-1, i * r + j)] = same(s, P.x_pos - 1, P.p_pos + i * r + j);
                g[P.NS(-2, i * r + j)] = same(s, P.x_pos - 2, P.p_pos + i * r + j);
            }
        }
    }
    {
        int i = (int)P.same_xm.size() - 2;
        Patch *s = P.ext_xm[i + 1];
        if (!P.same_xm[i + 1]) {
            std::vector<double> t = s->GetWenoValueFromCoarseLevel(P.x_pos - 1, P.p_pos + i * r, 3, val);
            for (int j = 0; j < 2; j++) {
                g[P.NS(-1, i * r + j)] = t[2 * j];
                g[P.NS(-2, i * r + j)] = t[2 * j + 1];
            }
        } else {
            for (int j = 0; j < 2; j++) {
                g[P.NS(-1, i * r + j)] = same(s, P.x_pos - 1, P.p_pos + i * r + j);
                g[P.NS(-2, i * r + j)] = same(s, P.x_pos - 2, P.p_pos + i * r + j);
            }
        }
    }
    {
        int i = -1;
        Patch *s = P.ext_xp[i + 1];
        if (!P.same_xp[i + 1]) {
            std::vector<double> t = s->GetWenoValueFromCoarseLevel(P.x_pos + P.n_x, P.p_pos + i * r, 1, val);
            for (int j = r - 2; j < r; j++) {
                g[P.NS(P.n_x, i * r + j)] = t[2 * j];
                g[P.NS(P.n_x + 1, i * r + j)] = t[2 * j + 1];
            }
        } else {
            for (int j = r - 2; j < r; j++) {
                g[P.NS(P.n_x, i * r + j)] = same(s, P.x_pos + P.n_x, P.p_pos + i * r + j);
                g[P.NS(P.n_x + 1, i * r + j)] = same(s, P.x_pos + P.n_x + 1, P.p_pos + i * r + j);
            }
        }
    }
    {
        int i = (int)P.same_xp.size() - 2;
        Patch *s = P.ext_xp[i + 1];
        if (!P.same_xp[i + 1]) {
            std::vector<double> t = s->GetWenoValueFromCoarseLevel(P.x_pos + P.n_x, P.p_pos + i * r, 1, val);
            for (int j = 0; j < 2; j++) {
                g[P.NS(P.n_x, i * r + j)] = t[2 * j];
                g[P.NS(P.n_x + 1, i * r + j)] = t[2 * j + 1];
            }
        } else {
            for (int j = 0; j < 2; j++) {
                g[P.NS(P.n_x, i * r + j)] = same(s, P.x_pos + P.n_x, P.p_pos + i * r + j);
                g[P.NS(P.n_x + 1, i * r + j)] = same(s, P.x_pos + P.n_x + 1, P.p_pos + i * r + j);
            }
        }
    }
}
}  // namespace

// Rectangle.cpp:417-560
void Patch::UpdateDifferentLevelBoundaries(int val) {
    std::vector<double> &g = f[val];
    for (unsigned i = 0; i < same_xm.size() - 2; i++)
        if (!same_xm[i + 1]) {
            std::vector<double> t = ext_xm[i + 1]->GetWenoValueFromCoarseLevel(x_pos - 1, p_pos + i * r, 3, val);
            for (int j = 0; j < r; j++) {
                g[NS(-1, i * r + j)] = t[2 * j];
                g[NS(-2, i * r + j)] = t[2 * j + 1];
            }
        }
    for (unsigned i = 0; i < same_xp.size() - 2; i++)
        if (!same_xp[i + 1]) {
            std::vector<double> t = ext_xp[i + 1]->GetWenoValueFromCoarseLevel(x_pos + n_x, p_pos + i * r, 1, val);
            for (int j = 0; j < r; j++) {
                g[NS(n_x, i * r + j)] = t[2 * j];
                g[NS(n_x + 1, i * r + j)] = t[2 * j + 1];
            }
        }
    for (unsigned i = 0; i < same_pm.size(); i++)
        if (!same_pm[i]) {
            std::vector<double> t = ext_pm[i]->GetWenoValueFromCoarseLevel(x_pos + i * r, p_pos - 1, 0, val);
            for (int j = 0; j < r; j++) {
                g[NS(i * r + j, -1)] = t[2 * j];
                g[NS(i * r + j, -2)] = t[2 * j + 1];
            }
        }
    for (unsigned i = 0; i < same_pp.size(); i++)
        if (!same_pp[i]) {
            std::vector<double> t = ext_pp[i]->GetWenoValueFromCoarseLevel(x_pos + i * r, p_pos + n_p, 2, val);
            for (int j = 0; j < r; j++) {
                g[NS(i * r + j, n_p)] = t[2 * j];
                g[NS(i * r + j, n_p + 1)] = t[2 * j + 1];
            }
        }
    corner_blocks(*this, val);
}

// Rectangle.cpp:1130-1214
void Patch::UpdateCornerPoints(int val) { corner_blocks(*this, val); }

// Rectangle.cpp:671-766
void Patch::CalculateConnectivitySame(Patch *rect) {
    int n_x_2 = rect->n_x, n_p_2 = rect->n_p, x_pos_2 = rect->x_pos, p_pos_2 = rect->p_pos;
    int x_pos_r = x_pos_2 - x_pos;
    int p_pos_r = p_pos_2 - p_pos;
    if (x_pos_r == n_x) {
        int lower = std::max(p_pos_2, p_pos);
        int upper = std::min(p_pos + n_p, p_pos_2 + n_p_2);
        if (OnLine(-1, p_pos_r, p_pos_r + n_p_2 - 1)) {
            ext_xp.at(0) = rect;
            same_xp.at(0) = 1;
        }
        if (OnLine(n_p, p_pos_r, p_pos_r + n_p_2 - 1)) {
            ext_xp[ext_xp.size() - 1] = rect;
            same_xp[ext_xp.size() - 1] = 1;
        }
        while (lower < upper) {
            int i1 = (lower - p_pos) / r + 1;
            ext_xp.at(i1) = rect;
            same_xp.at(i1) = 1;
            lower += r;
        }
    }
    if (x_pos_r == -n_x_2) {
        int lower = std::max(p_pos_2, p_pos);
        int upper = std::min(p_pos + n_p, p_pos_2 + n_p_2);
        if (OnLine(-1, p_pos_r, p_pos_r + n_p_2 - 1)) {
            ext_xm[0] = rect;
            same_xm[0] = 1;
        }
        if (OnLine(n_p, p_pos_r, p_pos_r + n_p_2 - 1)) {
            ext_xm[ext_xm.size() - 1] = rect;
            same_xm[ext_xm.size() - 1] = 1;
        }
        while (lower < upper) {
            int i1 = (lower - p_pos) / r + 1;
            ext_xm.at(i1) = rect;
            same_xm.at(i1) = 1;
            lower += r;
        }
    }
    if (p_pos_r == n_p) {
        int lower = std::max(x_pos_2, x_pos);
        int upper = std::min(x_pos + n_x, x_pos_2 + n_x_2);
        if (OnLine(-1, x_pos_r, x_pos_r + n_x_2 - 1)) {
            ext_xm[ext_xm.size() - 1] = rect;
            same_xm[ext_xm.size() - 1] = 1;
        }
        if (OnLine(n_x, x_pos_r, x_pos_r + n_x_2 - 1)) {
            ext_xp[ext_xp.size() - 1] = rect;
            same_xp[ext_xp.size() - 1] = 1;
        }
        while (lower < upper) {
            int i1 = (lower - x_pos) / r;
            ext_pp.at(i1) = rect;
            same_pp.at(i1) = 1;
            lower += r;
        }
    }
    if (p_pos_r == -n_p_2) {
        int lower = std::max(x_pos_2, x_pos);
        int upper = std::min(x_pos + n_x, x_pos_2 + n_x_2);
        if (OnLine(-1, x_pos_r, x_pos_r + n_x_2 - 1)) {
            ext_xm[0] = rect;
            same_xm[0] = 1;
        }
        if (OnLine(n_x, x_pos_r, x_pos_r + n_x_2 - 1)) {
            ext_xp[0] = rect;
            same_xp[0] = 1;
        }
        while (lower < upper) {
            int i1 = (lower - x_pos) / r;
            ext_pm.at(i1) = rect;
            same_pm.at(i1) = 1;
            lower += r;
        }
    }
}

// Rectangle.cpp:768-864 (this = coarse patch `current`, rect = finer patch)
void Patch::CalculateConnectivityFromFiner(Patch *current, Patch *rect) {
    int x_pos_2 = rect->x_pos / r - x_pos;
    int p_pos_2 = rect->p_pos / r - p_pos;
    int n_x_2 = rect->n_x / r;
    int n_p_2 = rect->n_p / r;
    int lower_x = std::max(0, x_pos_2);
    int upper_x = std::min(n_x, x_pos_2 + n_x_2);
    int lower_p = std::max(0, p_pos_2);
    int upper_p = std::min(n_p, p_pos_2 + n_p_2);
    for (int i = lower_x; i < upper_x; i++)
        for (int j = lower_p; j < upper_p; j++) {
            is_nested[NS(i, j)] = 1;
            finer_level[NS(i, j)] = rect;
        }
    if (((x_pos_2 + n_x_2) < (n_x + 1)) && ((x_pos_2 + n_x_2) > -1))
        for (int i = lower_p; i < upper_p; i++) {
            finer_level_x[NS(x_pos_2 + n_x_2, i)] = rect;
            is_ilb_x[NS(x_pos_2 + n_x_2, i)] = 1;
        }
    if (((x_pos_2 + n_x_2) < n_x) && ((x_pos_2 + n_x_2) > -1)) {
        for (int i = lower_p; i < upper_p; i++) {
            int j = i - p_pos_2 + 1;
            rect->ext_xp[j] = current;
            rect->same_xp[j] = 0;
        }
        if (OnLine(p_pos_2 - 1, 0, n_p - 1)) {
            rect->ext_xp[0] = current;
            rect->same_xp[0] = 0;
        }
        if (OnLine(p_pos_2 + n_p_2, 0, n_p - 1)) {
            rect->ext_xp[rect->ext_xp.size() - 1] = current;
            rect->same_xp[rect->ext_xp.size() - 1] = 0;
        }
    }
    if ((x_pos_2 < (n_x + 1)) && (x_pos_2 > -1))
        for (int i = lower_p; i < upper_p; i++) {
            finer_level_x[NS(x_pos_2, i)] = rect;
            is_ilb_x[NS(x_pos_2, i)] = 1;
        }
    if ((x_pos_2 < (n_x + 1)) && (x_pos_2 > 0)) {
        for (int i = lower_p; i < upper_p; i++) {
            int j = i - p_pos_2 + 1;
            rect->ext_xm[j] = current;
            rect->same_xm[j] = 0;
        }
        if (OnLine(p_pos_2 - 1, 0, n_p - 1)) {
            rect->ext_xm[0] = current;
            rect->same_xm[0] = 0;
        }
        if (OnLine(p_pos_2 + n_p_2, 0, n_p - 1)) {
            rect->ext_xm[rect->ext_xm.size() - 1] = current;
            rect->same_xm[rect->ext_xm.size() - 1] = 0;
        }
    }
    if (((p_pos_2 + n_p_2) < (n_p + 1)) && ((p_pos_2 + n_p_2) > -1))
        for (int i = lower_x; i < upper_x; i++) {
            finer_level_p[NS(i, p_pos_2 + n_p_2)] = rect;
            is_ilb_p[NS(i, p_pos_2 + n_p_2)] = 1;
        }
    if (((p_pos_2 + n_p_2) < n_p) && ((p_pos_2 + n_p_2) > -1))
        for (int i = lower_x; i < upper_x; i++) {
            int j = i - x_pos_2;
            rect->ext_pp[j] = current;
            rect->same_pp[j] = 0;
        }
    if ((p_pos_2 < (n_p + 1)) && (p_pos_2 > -1))
        for (int i = lower_x; i < upper_x; i++) {
            finer_level_p[NS(i, p_pos_2)] = rect;
            is_ilb_p[NS(i, p_pos_2)] = 1;
        }
    if ((p_pos_2 < (n_p + 1)) && (p_pos_2 > 0))
        for (int i = lower_x; i < upper_x; i++) {
            int j = i - x_pos_2;
            rect->ext_pm[j] = current;
            rect->same_pm[j] = 0;
        }
}

// Rectangle.hpp:128-130
double Patch::ErrorEstimate(int i, int j) const {
    const std::vector<double> &g = f[1];
    return errorWeights[0] * std::fabs(g[NS(i + 1, j)] - g[NS(i - 1, j)]) +
           errorWeights[1] * std::fabs(g[NS(i, j + 1)] - g[NS(i, j - 1)]) +
           errorWeights[2] * std::fabs(g[NS(i + 1, j)] - 2 * g[NS(i, j)] + g[NS(i - 1, j)]) +
           errorWeights[3] * std::fabs(g[NS(i, j + 1)] - 2 * g[NS(i, j)] + g[NS(i, j - 1)]) +
           errorWeights[4] * std::fabs(g[NS(i, j)]);
}

// Rectangle.cpp:866-890
void Patch::getError(std::vector<coords> &flagged) const {
    int offset = 0;
    for (int i = 0; i <= settings->maxDepth - depth; i++) offset += std::pow(settings->refinementRatio, i);
    offset *= 3;
    int lower_x = std::max(x_pos, offset);
    int upper_x = std::min(x_pos + n_x, settings->GetXSize(depth) - offset);
    int lower_p = std::max(p_pos, offset);
    int upper_p = std::min(p_pos + n_p, settings->GetPSize(depth, sp) - offset);
    for (int i = lower_x; i < upper_x; i++)
        for (int j = lower_p; j < upper_p; j++)
            if ((ErrorEstimate(i - x_pos, j - p_pos) > settings->refinementCriteria) ||
                settings->RefinementOverride((i + 0.5) * dx, Momentum(j - p_pos), depth, sp))
                flagged.push_back(std::make_pair(i, j));
}

// Rectangle.cpp:892-918 (this = old coarser patch, rect = new patch)
void Patch::GetDataFromCoarseLevelRectangle(Patch *rect) const {
    int x_pos_2 = rect->x_pos / r, p_pos_2 = rect->p_pos / r;
    int n_x_2 = rect->n_x / r, n_p_2 = rect->n_p / r;
    int lower_x = std::max(x_pos, x_pos_2 - 1);
    int upper_x = std::min(x_pos + n_x, x_pos_2 + n_x_2 + 1);
    int lower_p = std::max(p_pos, p_pos_2 - 1);
    int upper_p = std::min(p_pos + n_p, p_pos_2 + n_p_2 + 1);
    for (int i = lower_x; i < upper_x; i++)
        for (int j = lower_p; j < upper_p; j++) {
            std::vector<double> in = GetWenoValueFromCoarseLevel(r * i + 1, r * j + 1, -1, 1);
            for (int k = 0; k < r; k++)
                for (int l = 0; l < r; l++) {
                    int fi = (i - x_pos_2) * r + k, fj = (j - p_pos_2) * r + l;
                    rect->f[0][rect->NS(fi, fj)] = in[l + r * k];
                    rect->f[1][rect->NS(fi, fj)] = rect->f[0][rect->NS(fi, fj)];
                    rect->is_interpolated[rect->NS(fi, fj)] = 1;
                }
        }
}

// Rectangle.cpp:920-941 (this = old same-level patch, rect = new patch)
void Patch::GetDataFromSameLevelRectangle(Patch *rect) const {
    int x_pos_2 = rect->x_pos, p_pos_2 = rect->p_pos, n_x_2 = rect->n_x, n_p_2 = rect->n_p;
    int lower_x = std::max(x_pos, x_pos_2 - 2);
    int upper_x = std::min(x_pos + n_x, x_pos_2 + n_x_2 + 2);
    int lower_p = std::max(p_pos, p_pos_2 - 2);
    int upper_p = std::min(p_pos + n_p, p_pos_2 + n_p_2 + 2);
    for (int i = lower_x; i < upper_x; i++)
        for (int j = lower_p; j < upper_p; j++) {
            int ii = rect->NS(i - x_pos_2, j - p_pos_2);
            rect->f[0][ii] = GetValueFromSameLevel(i, j, 1);
            rect->f[1][ii] = rect->f[0][ii];
            rect->is_interpolated[ii] = 1;
        }
}

// Rectangle.cpp:1100-1128 (this = new coarser patch, rect = new finer patch)
void Patch::GetDataFromCoarseNewLevelRectangle(Patch *rect) const {
    int x_pos_2 = rect->x_pos / r, p_pos_2 = rect->p_pos / r;
    int n_x_2 = rect->n_x / r, n_p_2 = rect->n_p / r;
    int lower_x = std::max(x_pos, x_pos_2 - 1);
    int upper_x = std::min(x_pos + n_x, x_pos_2 + n_x_2 + 1);
    int lower_p = std::max(p_pos, p_pos_2 - 1);
    int upper_p = std::min(p_pos + n_p, p_pos_2 + n_p_2 + 1);
    for (int i = lower_x; i < upper_x; i++)
        for (int j = lower_p; j < upper_p; j++) {
            if (!rect->is_interpolated[rect->NS(r * (i - x_pos_2) + 1, r * (j - p_pos_2) + 1)]) {
                std::vector<double> in = GetWenoValueFromCoarseLevel(r * i + 1, r * j + 1, -1, 1);
                for (int k = 0; k < r; k++)
                    for (int l = 0; l < r; l++) {
                        int ii = rect->NS((i - x_pos_2) * r + k, (j - p_pos_2) * r + l);
                        rect->f[0][ii] = in[l + r * k];
                        rect->f[1][ii] = rect->f[0][ii];
                    }
            }
        }
}

// Rectangle.cpp:1032-1053
double Patch::RGKGetFluxX(int i, int j) const {
    if (!is_ilb_x[NS(i, j)]) {
        double w3 = 1.0 / 48.0;
        double dx_inv = 1 / dx;
        double dp_inv = 1 / dp;
        double q = settings->q.at(sp);
        double cc = kC * kC * settings->m.at(sp);
        double as = q * q * em->GetASquared(GetFinestIndex(i));
        double am = dp_inv * cc * (Gamma(Momentum(j + 1), as) - Gamma(Momentum(j), as));
        double ap1 = dp_inv * cc * (Gamma(Momentum(j + 2.0), as) - Gamma(Momentum(j + 1.0), as));
        double am1 = dp_inv * cc * (Gamma(Momentum(j), as) - Gamma(Momentum(j - 1.0), as));
        const std::vector<double> &g = f[1];
        double fm = weno_nomax(g[NS(i - 2, j)], g[NS(i - 1, j)], g[NS(i, j)], g[NS(i + 1, j)], am > 0.0);
        double fp1 = weno_nomax(g[NS(i - 2, j + 1)], g[NS(i - 1, j + 1)], g[NS(i, j + 1)], g[NS(i + 1, j + 1)], ap1 > 0.0);
        double fm1 = weno_nomax(g[NS(i - 2, j - 1)], g[NS(i - 1, j - 1)], g[NS(i, j - 1)], g[NS(i + 1, j - 1)], am1 > 0.0);
        return dx_inv * (fm * am + w3 * (fp1 - fm1) * (ap1 - am1));
    }
    return finer_level_x[NS(i, j)]->CalculateFluxToCoarseX(x_pos + i, p_pos + j);
}

// Rectangle.cpp:1069-1098
double Patch::RGKGetFluxP(int i, int j) const {
    if (!is_ilb_p[NS(i, j)]) {
        double w3 = 1 / 48.0;
        double dp_inv = 1 / dp;
        double dx_inv = 1 / dx;
        double cc = kC * kC * settings->m.at(sp);
        double q = settings->q.at(sp);
        double as_0 = q * q * em->GetASquared(GetFinestIndex(i - 1));
        double as_1 = q * q * em->GetASquared(GetFinestIndex(i));
        double as_2 = q * q * em->GetASquared(GetFinestIndex(i + 1));
        double as_3 = q * q * em->GetASquared(GetFinestIndex(i + 2));
        double Em = q * GetEfield(i);
        double Ep1 = q * GetEfield(i + 1);
        double Em1 = q * GetEfield(i - 1);
        double momentum = Momentum(j);
        double am = Em - cc * dx_inv * (Gamma(momentum, as_2) - Gamma(momentum, as_1));
        double ap1 = Ep1 - cc * dx_inv * (Gamma(momentum, as_3) - Gamma(momentum, as_2));
        double am1 = Em1 - cc * dx_inv * (Gamma(momentum, as_1) - Gamma(momentum, as_0));
        const std::vector<double> &g = f[1];
        double fm = weno_nomax(g[NS(i, j - 2)], g[NS(i, j - 1)], g[NS(i, j)], g[NS(i, j + 1)], am > 0.0);
        double fp1 = weno_nomax(g[NS(i + 1, j - 2)], g[NS(i + 1, j - 1)], g[NS(i + 1, j)], g[NS(i + 1, j + 1)], ap1 > 0.0);
        double fm1 = weno_nomax(g[NS(i - 1, j - 2)], g[NS(i - 1, j - 1)], g[NS(i - 1, j)], g[NS(i - 1, j + 1)], am1 > 0.0);
        return dp_inv * (fm * am + w3 * (fp1 - fm1) * (ap1 - am1));
    }
    return finer_level_p[NS(i, j)]->CalculateFluxToCoarseP(x_pos + i, p_pos + j);
}

// Rectangle.cpp:1216-1235
double Patch::RGKGetFluxPL(int i, int j) const {
    if (!is_ilb_p[NS(i, j)]) {
        double dp_inv = 1 / dp;
        double dx_inv = 1 / dx;
        double cc = kC * kC * settings->m.at(sp);
        double q = settings->q.at(sp);
        double as_1 = q * q * em->GetASquared(GetFinestIndex(i));
        double as_2 = q * q * em->GetASquared(GetFinestIndex(i + 1));
        double Em = q * GetEfield(i);
        double momentum = Momentum(j);
        double am = Em - cc * dx_inv * (Gamma(momentum, as_2) - Gamma(momentum, as_1));
        return dp_inv * (am > 0.0 ? f[1][NS(i, j - 1)] : f[1][NS(i, j)]) * am;
    }
    return finer_level_p[NS(i, j)]->CalculateFluxToCoarsePL(x_pos + i, p_pos + j);
}

// Rectangle.cpp:1237-1253
double Patch::RGKGetFluxXL(int i, int j) const {
    if (!is_ilb_x[NS(i, j)]) {
        double dx_inv = 1 / dx;
        double dp_inv = 1 / dp;
        double cc = kC * kC * settings->m.at(sp);
        double q = settings->q.at(sp);
        double as = q * q * em->GetASquared(GetFinestIndex(i));
        double am = dp_inv * cc * (Gamma(Momentum(j + 1), as) - Gamma(Momentum(j), as));
        return dx_inv * (am > 0.0 ? f[1][NS(i - 1, j)] : f[1][NS(i, j)]) * am;
    }
    return finer_level_x[NS(i, j)]->CalculateFluxToCoarseXL(x_pos + i, p_pos + j);
}

// Rectangle.hpp:132-178: coarse face flux = (1/r^2) * sum of the r fine fluxes
double Patch::CalculateFluxToCoarseP(int i, int j) const {
    i = i * r - x_pos;
    j = j * r - p_pos;
    double temp = 0.0;
    for (int k = 0; k < r; k++) temp += RGKGetFluxP(i + k, j);
    temp *= (1.0 / (r * r));
    return temp;
}
double Patch::CalculateFluxToCoarseX(int i, int j) const {
    i = i * r - x_pos;
    j = j * r - p_pos;
    double temp = 0.0;
    for (int k = 0; k < r; k++) temp += RGKGetFluxX(i, j + k);
    temp *= (1.0 / (r * r));
    return temp;
}
double Patch::CalculateFluxToCoarsePL(int i, int j) const {
    i = i * r - x_pos;
    j = j * r - p_pos;
    double temp = 0.0;
    for (int k = 0; k < r; k++) temp += RGKGetFluxPL(i + k, j);
    temp *= (1.0 / (r * r));
    return temp;
}
double Patch::CalculateFluxToCoarseXL(int i, int j) const {
    i = i * r - x_pos;
    j = j * r - p_pos;
    double temp = 0.0;
    for (int k = 0; k < r; k++) temp += RGKGetFluxXL(i, j + k);
    temp *= (1.0 / (r * r));
    return temp;
}

// Rectangle.cpp:1625-1706 (this = coarse patch, rect = finer caller)
void Patch::SetCFromDifferentLevel(int i, int j, Patch *rect, int t) {
    int i_caller = i - rect->x_pos;
    int j_caller = j - rect->p_pos;
    int i_coarse = i / r - x_pos;
    int j_coarse = j / r - p_pos;
    Patch &R = *rect;
    if (t == 0) {
        double cx = Cx[NS(i_coarse, j_coarse)];
        for (int k = 0; k < r; k++)
            cx = std::min(cx, R.FxDS[R.NS(i_caller, j_caller + k)] > 0.0 ? R.Rp[R.NS(i_caller, j_caller + k)]
                                                                       : R.Rm[R.NS(i_caller, j_caller + k)]);
        cx = std::min(cx, FxDS[NS(i_coarse, j_coarse)] > 0.0 ? Rm[NS(i_coarse - 1, j_coarse)]
                                                             : Rp[NS(i_coarse - 1, j_coarse)]);
        for (int k = 0; k < r; k++) R.Cx[R.NS(i_caller, j_caller + k)] = cx;
        Cx[NS(i_coarse, j_coarse)] = cx;
    } else if (t == 1) {
        double cx = Cx[NS(i_coarse, j_coarse)];
        for (int k = 0; k < r; k++)
            cx = std::min(cx, R.FxDS[R.NS(i_caller, j_caller + k)] > 0.0 ? R.Rm[R.NS(i_caller - 1, j_caller + k)]
                                                                       : R.Rp[R.NS(i_caller - 1, j_caller + k)]);
        cx = std::min(cx, FxDS[NS(i_coarse, j_coarse)] > 0.0 ? Rp[NS(i_coarse, j_coarse)]
                                                             : Rm[NS(i_coarse, j_coarse)]);
        for (int k = 0; k < r; k++) R.Cx[R.NS(i_caller, j_caller + k)] = cx;
        Cx[NS(i_coarse, j_coarse)] = cx;
    } else if (t == 2) {
        double cp = Cp[NS(i_coarse, j_coarse)];
        for (int k = 0; k < r; k++)
            cp = std::min(cp, R.FpDS[R.NS(i_caller + k, j_caller)] > 0.0 ? R.Rp[R.NS(i_caller + k, j_caller)]
                                                                       : R.Rm[R.NS(i_caller + k, j_caller)]);
        cp = std::min(cp, FpDS[NS(i_coarse, j_coarse)] > 0.0 ? Rm[NS(i_coarse, j_coarse - 1)]
                                                             : Rp[NS(i_coarse, j_coarse - 1)]);
        for (int k = 0; k < r; k++) R.Cp[R.NS(i_caller + k, j_caller)] = cp;
        Cp[NS(i_coarse, j_coarse)] = cp;
    } else if (t == 3) {
        double cp = Cp[NS(i_coarse, j_coarse)];
        for (int k = 0; k < r; k++)
            cp = std::min(cp, R.FpDS[R.NS(i_caller + k, j_caller)] > 0.0 ? R.Rm[R.NS(i_caller + k, j_caller - 1)]
                                                                       : R.Rp[R.NS(i_caller + k, j_caller - 1)]);
        cp = std::min(cp, FpDS[NS(i_coarse, j_coarse)] > 0.0 ? Rp[NS(i_coarse, j_coarse)]
                                                             : Rm[NS(i_coarse, j_coarse)]);
        for (int k = 0; k < r; k++) R.Cp[R.NS(i_caller + k, j_caller)] = cp;
        Cp[NS(i_coarse, j_coarse)] = cp;
    }
}

// Rectangle.cpp:1708-1760
void Patch::CalculateDifferentBoundaryC() {
    for (unsigned i = 0; i < same_xm.size() - 2; i++)
        if (!same_xm[i + 1]) ext_xm[i + 1]->SetCFromDifferentLevel(x_pos, p_pos + r * i, this, 0);
    for (unsigned i = 0; i < same_xp.size() - 2; i++)
        if (!same_xp[i + 1]) ext_xp[i + 1]->SetCFromDifferentLevel(x_pos + n_x, p_pos + r * i, this, 1);
    for (unsigned i = 0; i < same_pm.size(); i++)
        if (!same_pm[i]) ext_pm[i]->SetCFromDifferentLevel(x_pos + r * i, p_pos, this, 2);
    for (unsigned i = 0; i < same_pp.size(); i++)
        if (!same_pp[i]) ext_pp[i]->SetCFromDifferentLevel(x_pos + r * i, p_pos + n_p, this, 3);
}

// Rectangle.cpp:1794-1830 (this = neighbour, rect = caller)
void Patch::SetCFromSameLevel(int i, int j, Patch *rect, int t) {
    int i_n = i - x_pos, j_n = j - p_pos;
    int i_c = i - rect->x_pos, j_c = j - rect->p_pos;
    Patch &R = *rect;
    if (t == 0) {
        for (int k = 0; k < r; k++) {
            double cx = FxDS[NS(i_n, j_n + k)] > 0.0 ? std::min(R.Rp[R.NS(i_c, j_c + k)], Rm[NS(i_n - 1, j_n + k)])
                                                     : std::min(Rp[NS(i_n - 1, j_n + k)], R.Rm[R.NS(i_c, j_c + k)]);
            Cx[NS(i_n, j_n + k)] = cx;
            R.Cx[R.NS(i_c, j_c + k)] = cx;
        }
    } else if (t == 1) {
        for (int k = 0; k < r; k++) {
            double cx = FxDS[NS(i_n, j_n + k)] > 0.0 ? std::min(R.Rm[R.NS(i_c - 1, j_c + k)], Rp[NS(i_n, j_n + k)])
                                                     : std::min(Rm[NS(i_n, j_n + k)], R.Rp[R.NS(i_c - 1, j_c + k)]);
            Cx[NS(i_n, j_n + k)] = cx;
            R.Cx[R.NS(i_c, j_c + k)] = cx;
        }
    } else if (t == 2) {
        for (int k = 0; k < r; k++) {
            double cp = FpDS[NS(i_n + k, j_n)] > 0.0 ? std::min(R.Rp[R.NS(i_c + k, j_c)], Rm[NS(i_n + k, j_n - 1)])
                                                     : std::min(Rp[NS(i_n + k, j_n - 1)], R.Rm[R.NS(i_c + k, j_c)]);
            Cp[NS(i_n + k, j_n)] = cp;
            R.Cp[R.NS(i_c + k, j_c)] = cp;
        }
    } else if (t == 3) {
        for (int k = 0; k < r; k++) {
            double cp = FpDS[NS(i_n + k, j_n)] > 0.0 ? std::min(R.Rm[R.NS(i_c + k, j_c - 1)], Rp[NS(i_n + k, j_n)])
                                                     : std::min(Rm[NS(i_n + k, j_n)], R.Rp[R.NS(i_c + k, j_c - 1)]);
            Cp[NS(i_n + k, j_n)] = cp;
            R.Cp[R.NS(i_c + k, j_c)] = cp;
        }
    }
}

// Rectangle.cpp:1832-1877.  A domain-edge neighbour is the sentinel, whose
// arrays are empty in the reference (the call is undefined behaviour there and
// could only touch wall faces, which no update loop reads): skipped.
void Patch::CalculateSameBoundaryC() {
    for (unsigned i = 0; i < same_xm.size() - 2; i++)
        if (same_xm[i + 1] && ext_xm[i + 1]) ext_xm[i + 1]->SetCFromSameLevel(x_pos, p_pos + i * r, this, 0);
    for (unsigned i = 0; i < same_xp.size() - 2; i++)
        if (same_xp[i + 1] && ext_xp[i + 1]) ext_xp[i + 1]->SetCFromSameLevel(x_pos + n_x, p_pos + i * r, this, 1);
    for (unsigned i = 0; i < same_pm.size(); i++)
        if (same_pm[i] && ext_pm[i]) ext_pm[i]->SetCFromSameLevel(x_pos + i * r, p_pos, this, 2);
    for (unsigned i = 0; i < same_pp.size(); i++)
        if (same_pp[i] && ext_pp[i]) ext_pp[i]->SetCFromSameLevel(x_pos + i * r, p_pos + n_p, this, 3);
}

// Rectangle.cpp:1879-1915
void Patch::UpdateCFromSameLevel(int i, int j, Patch *rect, int t) {
    int i_n = i - x_pos, j_n = j - p_pos;
    int i_c = i - rect->x_pos, j_c = j - rect->p_pos;
    Patch &R = *rect;
    if (t == 0 || t == 1) {
        for (int k = 0; k < r; k++) {
            double cx = std::min(Cx[NS(i_n, j_n + k)], R.Cx[R.NS(i_c, j_c + k)]);
            Cx[NS(i_n, j_n + k)] = cx;
            R.Cx[R.NS(i_c, j_c + k)] = cx;
        }
    } else if (t == 2 || t == 3) {
        for (int k = 0; k < r; k++) {
            double cp = std::min(Cp[NS(i_n + k, j_n)], R.Cp[R.NS(i_c + k, j_c)]);
            Cp[NS(i_n + k, j_n)] = cp;
            R.Cp[R.NS(i_c + k, j_c)] = cp;
        }
    }
}

// Rectangle.cpp:1917-1960 (sentinel neighbours skipped as above)
void Patch::UpdateSameBoundaryC() {
    for (unsigned i = 0; i < same_xm.size() - 2; i++)
        if (same_xm[i + 1] && ext_xm[i + 1]) ext_xm[i + 1]->UpdateCFromSameLevel(x_pos, p_pos + i * r, this, 0);
    for (unsigned i = 0; i < same_xp.size() - 2; i++)
        if (same_xp[i + 1] && ext_xp[i + 1]) ext_xp[i + 1]->UpdateCFromSameLevel(x_pos + n_x, p_pos + i * r, this, 1);
    for (unsigned i = 0; i < same_pm.size(); i++)
        if (same_pm[i] && ext_pm[i]) ext_pm[i]->UpdateCFromSameLevel(x_pos + i * r, p_pos, this, 2);
    for (unsigned i = 0; i < same_pp.size(); i++)
        if (same_pp[i] && ext_pp[i]) ext_pp[i]->UpdateCFromSameLevel(x_pos + i * r, p_pos + n_p, this, 3);
}

// ---------------------------------------------------------------- Level
// Level.cpp:128-150
void Level::GetDataFromSameLevel(Level &lvl) {
    for (auto &old : lvl.rectangles)
        for (auto &nw : rectangles) old->GetDataFromSameLevelRectangle(nw.get());
}
void Level::GetDataFromCoarserLevel(Level &lvl) {
    for (auto &old : lvl.rectangles)
        for (auto &nw : rectangles) old->GetDataFromCoarseLevelRectangle(nw.get());
}
void Level::GetDataFromCoarseNewLevel(Level &lvl) {
    for (auto &coarse : lvl.rectangles)
        for (auto &fine : rectangles) coarse->GetDataFromCoarseNewLevelRectangle(fine.get());
}

// ---------------------------------------------------------------- Mesh

// Mesh.cpp:8-50
Mesh::Mesh(int sp_, Settings &s) : settings(s), sp(sp_) {
    std::vector<coords> flaggedCells;
    rect extrema;
    level identified;
    for (int l = 0; l <= s.maxDepth; l++) {
        if (l > 0) {
            if (flaggedCells.size() > 0) {
                getExtrema(extrema, flaggedCells);
                identified = splitRectangle(extrema, flaggedCells, s.minEfficiency);
                interpRectanglesUp(identified, l - 1);
                hierarchy.push_back(identified);
            }
            if (l <= s.maxDepth) {
                flaggedCells.clear();
                getError(l, true, flaggedCells);
            }
        } else {
            int xmax = s.GetXSize(s.maxDepth), pmax = s.GetPSize(s.maxDepth, sp);
            extrema = std::make_pair(std::make_pair(0, 0), std::make_pair(xmax, pmax));
            identified.push_back(extrema);
            hierarchy.push_back(identified);
            getError(l, true, flaggedCells);
        }
    }
    promoteHierarchyToMesh(true);
    for (int i = 0; i < s.maxDepth; i++) {
        PushData();
        updateHierarchy(true);
    }
}

// Mesh.cpp:132-210
void Mesh::updateHierarchy(bool init) {
    std::vector<coords> flaggedCells, flaggedl2, flaggedl1, flaggedtmp;
    rect extrema;
    level identified;
    if (hierarchy.empty()) throw std::runtime_error("hierarchy is empty, cannot update it");
    int lloop = (int)hierarchy.size() - 1;
    for (int l = lloop; l > -1; l--) {
        if (l < (int)hierarchy.size() - 2) {
            flaggedl1.clear();
            flaggedl2.clear();
            if (hierarchy.size() > 2) {
                for (rect &rr : hierarchy.at(l + 2)) {
                    flaggedtmp.clear();
                    mergeDownFlaggedData(l, rr, flaggedtmp);
                    if (!flaggedtmp.empty())
                        flaggedl2.insert(flaggedl2.end(), flaggedtmp.begin(), flaggedtmp.end());
                }
            }
            getError(l, false, flaggedl1);
            std::sort(flaggedl1.begin(), flaggedl1.end());
            flaggedCells.reserve(flaggedl1.size() + flaggedl2.size());
            std::set_union(flaggedl1.begin(), flaggedl1.end(), flaggedl2.begin(), flaggedl2.end(),
                           std::back_inserter(flaggedCells));
        } else {
            getError(l, false, flaggedCells);
        }
        if (flaggedCells.size() > 0) {
            getExtrema(extrema, flaggedCells);
            identified = splitRectangle(extrema, flaggedCells, settings.minEfficiency);
            if (l == lloop) {
                if (l < settings.maxDepth) {
                    interpRectanglesUp(identified, l);
                    hierarchy.push_back(identified);
                }
            } else {
                interpRectanglesUp(identified, l);
                hierarchy.at(l + 1) = identified;
            }
        } else {
            if (l < lloop) hierarchy.erase(hierarchy.begin() + l + 1);
        }
        flaggedCells.clear();
    }
    promoteHierarchyToMesh(init);
}

// Mesh.cpp:212-295.  With MKLINIT == 0 (veritas.hpp:33) the initial branch
// collects its flags in a private vector that is never merged, so it flags
// nothing; the time-advance branch asks each patch of the level.
void Mesh::getError(int lvl, bool init, std::vector<coords> &flagged) {
    if (init) return;
    for (auto &rr : levels.at(settings.maxDepth - lvl)->rectangles) rr->getError(flagged);
}

// Mesh.cpp:298-313
void Mesh::interpRectanglesUp(level &identified, int lvl) {
    int nmax = settings.GetXSize(settings.maxDepth - lvl);
    int pmax = settings.GetPSize(settings.maxDepth - lvl, sp);
    int nmax1 = settings.GetXSize(settings.maxDepth - (lvl + 1));
    int pmax1 = settings.GetPSize(settings.maxDepth - (lvl + 1), sp);
    for (rect &rr : identified) {
        rr.second.first += 1;
        rr.second.second += 1;
        rr.first.first *= (nmax1 / double(nmax));
        rr.first.second *= (pmax1 / double(pmax));
        rr.second.first *= (nmax1 / double(nmax));
        rr.second.second *= (pmax1 / double(pmax));
    }
}

// Mesh.cpp:315-339
void Mesh::mergeDownFlaggedData(int lvl, const rect &rr, std::vector<coords> &found) {
    int ic, jc;
    int nmax = settings.GetXSize(settings.maxDepth - lvl);
    int pmax = settings.GetPSize(settings.maxDepth - lvl, sp);
    int nmaxl2 = settings.GetXSize(settings.maxDepth - (lvl + 2));
    int pmaxl2 = settings.GetPSize(settings.maxDepth - (lvl + 2), sp);
    int R = (int)settings.refinementRatio;
    for (int i = rr.first.first - 2 * R; i <= rr.second.first + 2 * R + 1; i++)
        for (int j = rr.first.second - 2 * R; j <= rr.second.second + 2 * R; j++) {
            ic = i * (nmax / double(nmaxl2));
            jc = j * (pmax / double(pmaxl2));
            if ((ic > -1) && (ic < nmax) && (jc > -1) && (jc < pmax)) found.push_back(std::make_pair(ic, jc));
        }
    std::sort(found.begin(), found.end());
    auto it = std::unique(found.begin(), found.end());
    found.resize(std::distance(found.begin(), it));
}

// Mesh.cpp:361-364
int Mesh::countCells(const rect &span) {
    return (span.second.first - span.first.first + 1) * (span.second.second - span.first.second + 1);
}

// Mesh.cpp:366-371
void Mesh::isFlaggedInside(const std::vector<coords> &flagged, std::vector<bool> &inside, const rect &rr) {
    for (unsigned i = 0; i < flagged.size(); i++)
        inside.at(i) = ((flagged.at(i).first >= rr.first.first) && (flagged.at(i).first <= rr.second.first) &&
                        (flagged.at(i).second >= rr.first.second) && (flagged.at(i).second <= rr.second.second)) == 1;
}

// Mesh.cpp:373-447 (signatures and their discrete Laplacians)
std::tuple<std::vector<int>, std::vector<int>, std::vector<int>, std::vector<int>>
Mesh::computeSignatures(const rect &rr, const std::vector<coords> &flagged) {
    double lengthx = rr.second.first - rr.first.first;
    double lengthp = rr.second.second - rr.first.second;
    std::vector<int> sigX(lengthx + 1), sigP(lengthp + 1);
    std::vector<int> delSigX, delSigP;
    for (const coords &xp : flagged) {
        sigX.at(xp.first - rr.first.first)++;
        sigP.at(xp.second - rr.first.second)++;
    }
    if (lengthx >= 2) {
        delSigX.resize(lengthx - 1);
        for (unsigned h = 1; h < sigX.size() - 1; h++)
            delSigX.at(h - 1) = sigX.at(h - 1) + sigX.at(h + 1) - 2 * sigX.at(h);
    }
    if (lengthp >= 2) {
        delSigP.resize(lengthp - 1);
        for (unsigned h = 1; h < sigP.size() - 1; h++)
            delSigP.at(h - 1) = sigP.at(h - 1) + sigP.at(h + 1) - 2 * sigP.at(h);
    }
    return std::make_tuple(sigX, delSigX, sigP, delSigP);
}

static int sgn(int x) { return (x > 0) ? 1 : ((x < 0) ? -1 : 0); }

// Mesh.cpp:449-595
coords Mesh::identifyInflection(
    const rect &rr,
    const std::tuple<std::vector<int>, std::vector<int>, std::vector<int>, std::vector<int>> &sig) {
    int inflectX = 0, inflectP = 0, inflect = 0, indexX = 0, indexP = 0, sizeSigX, sizeSigP, mx, countmax = 0;
    std::vector<int> inflectIdx, inflectSize;
    std::vector<int>::iterator maxit;
    std::pair<bool, bool> bisect = std::make_pair(false, false);
    const std::vector<int> &delSigX = std::get<1>(sig);
    const std::vector<int> &delSigP = std::get<3>(sig);
    sizeSigX = delSigX.size();
    sizeSigP = delSigP.size();
    if (sizeSigX > 2) {
        for (unsigned i = 0; i < delSigX.size() - 1; i++)
            if ((delSigX.at(i) != 0) && (delSigX.at(i + 1) != 0) && (sgn(delSigX.at(i)) != sgn(delSigX.at(i + 1)))) {
                inflectIdx.push_back(i);
                inflectSize.push_back(std::abs(delSigX.at(i + 1) - delSigX.at(i)));
            }
        inflect = inflectIdx.size();
        if (!inflectSize.empty()) {
            maxit = std::max_element(inflectSize.begin(), inflectSize.end());
            mx = *maxit;
            countmax = std::count(inflectSize.begin(), inflectSize.end(), mx);
        }
        if (inflect == 1) {
            indexX = inflectIdx.at(0);
            inflectX = inflectSize.at(0);
        } else if (inflect > 1) {
            if (countmax > 1) {
                int center = sizeSigX / 2;
                int range = std::numeric_limits<int>::max();
                int chosen = 0;
                for (unsigned i = 0; i < inflectIdx.size(); i++)
                    if (range > std::abs(inflectIdx.at(i) - center)) {
                        range = std::abs(inflectIdx.at(i) - center);
                        chosen = i;
                    }
                indexX = inflectIdx.at(chosen);
                inflectX = inflectSize.at(chosen);
            } else {
                indexX = inflectIdx.at(maxit - inflectSize.begin());
                inflectX = inflectSize.at(maxit - inflectSize.begin());
            }
        } else {
            bisect.first = true;
        }
    }
    inflectIdx.clear();
    inflectSize.clear();
    if (sizeSigP > 2) {
        for (unsigned i = 0; i < delSigP.size() - 1; i++)
            if ((delSigP.at(i) != 0) && (delSigP.at(i + 1) != 0) && (sgn(delSigP.at(i)) != sgn(delSigP.at(i + 1)))) {
                inflectIdx.push_back(i);
                inflectSize.push_back(std::abs(delSigP.at(i + 1) - delSigP.at(i)));
            }
        inflect = inflectIdx.size();
        if (!inflectSize.empty()) {
            maxit = std::max_element(inflectSize.begin(), inflectSize.end());
            mx = *maxit;
            countmax = std::count(inflectSize.begin(), inflectSize.end(), mx);
        }
        if (inflect == 1) {
            indexP = inflectIdx.at(0);
            inflectP = inflectSize.at(0);
        } else if (inflect > 1) {
            if (countmax > 1) {
                int center = sizeSigP / 2;
                int range = 32767;
                int chosen = 0;
                for (unsigned i = 0; i < inflectIdx.size(); i++)
                    if (range > std::abs(inflectIdx.at(i) - center)) {
                        range = std::abs(inflectIdx.at(i) - center);
                        chosen = i;
                    }
                indexP = inflectIdx.at(chosen);
                inflectP = inflectSize.at(chosen);
            } else {
                indexP = inflectIdx.at(maxit - inflectSize.begin());
                inflectP = inflectSize.at(maxit - inflectSize.begin());
            }
        } else {
            bisect.second = true;
        }
    }
    if (bisect.first && bisect.second) return std::make_pair(2, 0);
    if (((sizeSigX <= 2) && (sizeSigP <= 2)) || ((inflectX == 0) && (inflectP == 0))) return std::make_pair(3, 0);
    if (inflectX > inflectP) {
        indexX += rr.first.first + 1;
        return std::make_pair(0, indexX);
    }
    indexP += rr.first.second + 1;
    return std::make_pair(1, indexP);
}

// Mesh.cpp:597-624
std::tuple<bool, int, int> Mesh::hasHole(const std::vector<int> &sig) {
    int start = 0, finish = 0;
    auto it = std::search_n(sig.begin(), sig.end(), 1, 0);
    bool found = it != sig.end();
    if (found) {
        it == sig.begin() ? start = it - sig.begin() : start = it - sig.begin() - 1;
        auto it2 = std::find_if_not(it, sig.end(), [](int i) { return i == 0; });
        it2 == sig.end() ? finish = it2 - sig.begin() - 1 : finish = it2 - sig.begin();
    }
    return std::make_tuple(found, start, finish);
}

// Mesh.cpp:626-642
void Mesh::getExtrema(rect &extrema, const std::vector<coords> &flagged) {
    auto byx = [](const coords &a, const coords &b) { return a.first < b.first; };
    auto byp = [](const coords &a, const coords &b) { return a.second < b.second; };
    auto limx = std::minmax_element(flagged.begin(), flagged.end(), byx);
    auto limp = std::minmax_element(flagged.begin(), flagged.end(), byp);
    coords ll = std::make_pair(limx.first->first, limp.first->second);
    coords ur = std::make_pair(limx.second->first, limp.second->second);
    extrema = std::make_pair(ll, ur);
}

// Mesh.cpp:644-792 (recursive bisection)
level Mesh::splitRectangle(const rect &rectangle, const std::vector<coords> &flagged, double minEfficiency) {
    std::vector<rect> resultRect;
    double efficiencyRatio = flagged.size() / double(countCells(rectangle));
    if (efficiencyRatio > minEfficiency) {
        resultRect.push_back(rectangle);
        return resultRect;
    }
    std::tuple<bool, int, int> holePosX, holePosP;
    rect pieceOne, pieceTwo;
    std::vector<coords> flaggedOne, flaggedTwo;
    std::vector<bool> isInside;
    std::vector<rect> resultRectTwo, resultMerged;
    auto signatures = computeSignatures(rectangle, flagged);
    holePosX = hasHole(std::get<0>(signatures));
    holePosP = hasHole(std::get<2>(signatures));
    std::get<1>(holePosX) += rectangle.first.first;
    std::get<2>(holePosX) += rectangle.first.first;
    std::get<1>(holePosP) += rectangle.first.second;
    std::get<2>(holePosP) += rectangle.first.second;
    if (std::get<0>(holePosX) && std::get<0>(holePosP)) {
        if (std::abs(std::get<2>(holePosX) - std::get<1>(holePosX)) > std::abs(std::get<2>(holePosP) - std::get<1>(holePosP))) {
            pieceOne = std::make_pair(rectangle.first, coords(std::get<1>(holePosX), rectangle.second.second));
            pieceTwo = std::make_pair(coords(std::get<2>(holePosX), rectangle.first.second), rectangle.second);
        } else {
            pieceOne = std::make_pair(rectangle.first, coords(rectangle.second.first, std::get<1>(holePosP)));
            pieceTwo = std::make_pair(coords(rectangle.first.first, std::get<2>(holePosP)), rectangle.second);
        }
    } else if (std::get<0>(holePosX)) {
        pieceOne = std::make_pair(rectangle.first, coords(std::get<1>(holePosX), rectangle.second.second));
        pieceTwo = std::make_pair(coords(std::get<2>(holePosX), rectangle.first.second), rectangle.second);
    } else if (std::get<0>(holePosP)) {
        pieceOne = std::make_pair(rectangle.first, coords(rectangle.second.first, std::get<1>(holePosP)));
        pieceTwo = std::make_pair(coords(rectangle.first.first, std::get<2>(holePosP)), rectangle.second);
    } else {
        coords split = identifyInflection(rectangle, signatures);
        if (split.first == 0) {
            pieceOne = std::make_pair(rectangle.first, coords(split.second, rectangle.second.second));
            pieceTwo = std::make_pair(coords(split.second + 1, rectangle.first.second), rectangle.second);
        } else if (split.first == 1) {
            pieceOne = std::make_pair(rectangle.first, coords(rectangle.second.first, split.second));
            pieceTwo = std::make_pair(coords(rectangle.first.first, split.second + 1), rectangle.second);
        } else if (split.first == 2) {
            double cutPosition, lengthx, lengthp;
            lengthx = rectangle.second.first - rectangle.first.first;
            lengthp = rectangle.second.second - rectangle.first.second;
            if (lengthx > lengthp) {
                cutPosition = std::floor((lengthx / 2) + rectangle.first.first);
                pieceOne = std::make_pair(rectangle.first, coords(cutPosition, rectangle.second.second));
                pieceTwo = std::make_pair(coords(cutPosition + 1, rectangle.first.second), rectangle.second);
            } else {
                cutPosition = std::floor((lengthp / 2) + rectangle.first.second);
                pieceOne = std::make_pair(rectangle.first, coords(rectangle.second.first, cutPosition));
                pieceTwo = std::make_pair(coords(rectangle.first.first, cutPosition + 1), rectangle.second);
            }
        } else {
            resultRect.push_back(rectangle);
            return resultRect;
        }
    }
    isInside.resize(flagged.size());
    isFlaggedInside(flagged, isInside, pieceOne);
    for (unsigned i = 0; i < flagged.size(); i++) {
        if (isInside.at(i))
            flaggedOne.push_back(flagged.at(i));
        else
            flaggedTwo.push_back(flagged.at(i));
    }
    if (flaggedOne.size() > 0) resultRect = splitRectangle(pieceOne, flaggedOne, minEfficiency);
    if (flaggedTwo.size() > 0) resultRectTwo = splitRectangle(pieceTwo, flaggedTwo, minEfficiency);
    resultMerged.reserve(resultRect.size() + resultRectTwo.size());
    resultMerged.insert(resultMerged.end(), resultRect.begin(), resultRect.end());
    resultMerged.insert(resultMerged.end(), resultRectTwo.begin(), resultRectTwo.end());
    return resultMerged;
}

// Mesh.cpp:794-875.  Note that a level's extent is taken at N - lvl, N + 1 the
// number of levels in use, so the up/right walls of a hierarchy with unused
// levels are compared with the size of a finer level (a result-affecting quirk
// kept as written).
void Mesh::promoteHierarchyToMesh(bool init) {
    int xmax, pmax, N = (int)hierarchy.size() - 1;
    int empty = settings.maxDepth - N;
    bool up, down, left, right;
    std::vector<std::unique_ptr<Level>> levels_n;
    if (!init)
        for (auto &lvl : levels) levels_n.push_back(std::move(lvl));
    levels.clear();
    for (int i = 0; i < empty; i++) levels.push_back(std::unique_ptr<Level>(new Level(sp, i, settings)));
    for (int lvl = N; lvl >= 0; lvl--) {
        xmax = settings.GetXSize(N - lvl);
        pmax = settings.GetPSize(N - lvl, sp);
        levels.push_back(std::unique_ptr<Level>(new Level(sp, N - lvl + empty, settings)));
        for (rect &rr : hierarchy.at(lvl)) {
            up = rr.second.second == pmax;
            down = rr.first.second == 0;
            left = rr.first.first == 0;
            right = rr.second.first == xmax;
            levels.at(N - lvl + empty)->rectangles.push_back(std::unique_ptr<Patch>(
                new Patch(rr.second.first - rr.first.first, rr.second.second - rr.first.second, rr.first.first,
                          rr.first.second, N - lvl + empty, settings, up, down, left, right, sp)));
        }
    }
    if (init)
        for (auto &lvl : levels)
            for (auto &rr : lvl->rectangles) rr->InitializeDistribution();
    for (unsigned l = 0; l < levels.size(); l++)
        for (auto &rl : levels.at(l)->rectangles)
            if (l > 0)
                for (auto &rf : levels.at(l - 1)->rectangles) rl->CalculateConnectivityFromFiner(rl.get(), rf.get());
    for (unsigned l = 0; l < levels.size(); l++)
        for (auto &rl : levels.at(l)->rectangles)
            for (auto &rr : levels.at(l)->rectangles)
                if (rl != rr) rl->CalculateConnectivitySame(rr.get());
    SetFieldSolver(em);
    if (!init) InterMeshDataTransfer(levels_n);
    PushData();
    for (unsigned l = 0; l < levels.size(); l++)
        for (auto &rl : levels.at(l)->rectangles) rl->FCTTimeStep(0, -1, 3);
}

// Mesh.cpp:116-130
void Mesh::InterMeshDataTransfer(std::vector<std::unique_ptr<Level>> &levels_n) {
    for (unsigned i = 0; i < levels_n.size() - 1; i++) {
        levels[i]->GetDataFromCoarserLevel(*levels_n[i + 1]);
        levels[i]->GetDataFromSameLevel(*levels_n[i]);
    }
    if (levels_n.size() > 1) {
        int i = (int)levels_n.size() - 1;
        levels[i]->GetDataFromSameLevel(*levels_n[i]);
    }
    for (int i = (int)levels.size() - 1; i > 0; i--) levels[i - 1]->GetDataFromCoarseNewLevel(*levels[i]);
}

// ---------------------------------------------------------------- Solver
// SolverManager.cpp:48-58 (without the screen output)
void Solver::reGrid() {
    for (auto &m : meshes) m->updateHierarchy();
}

}  // namespace orc
