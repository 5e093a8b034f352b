// CPU ORACLE C API -- test infrastructure only.  Loaded by tests/ (ctypes),
// __graft_entry__.smoke() and bench.py's CPU-baseline leg.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.hpp"
#include "../veritas_amd/csrc/vx_math.h"

namespace {
struct Run {
    vx::Scenario scn;
    std::unique_ptr<orc::Solver> s;
};

orc::Patch &base(Run *r) { return *r->s->meshes[0]->levels.back()->rectangles[0]; }

int copy_out(const std::vector<double> &v, double *out, int n) {
    int m = (int)v.size();
    if (out && n >= m) std::memcpy(out, v.data(), sizeof(double) * m);
    return m;
}

// interior of a patch plane, row-major (i over n_x, j over n_p)
int copy_plane(const orc::Patch &p, const std::vector<double> &v, int ghosts, double *out, int n) {
    int g = ghosts;
    int nx = p.n_x + 2 * g, np = p.n_p + 2 * g;
    if (out && n >= nx * np)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < np; j++) out[i * np + j] = v[p.NS(i - g, j - g)];
    return nx * np;
}
}  // namespace

extern "C" {

void *vxo_create(const vx_scenario_params *prm) {
    Run *r = new Run;
    r->scn = vx::Scenario::from_params(*prm);
    r->s.reset(new orc::Solver(r->scn.desc(), r->scn.hooks()));
    return r;
}

// A run of the reference's inputs and user hooks (`hooks->user` must stay
// valid while the run lives); nullptr when the inputs are rejected.
void *vxo_create_desc(const vx_settings_desc *desc, const vx_hooks *hooks) {
    Run *r = new Run;
    try {
        r->s.reset(new orc::Solver(*desc, *hooks));
    } catch (std::exception &) {
        delete r;
        return nullptr;
    }
    return r;
}

void vxo_destroy(void *h) { delete (Run *)h; }

double vxo_calc_dt(void *h, double cfl) { return ((Run *)h)->s->CalculateDt(cfl); }

void vxo_advance(void *h, double dt, int nsteps) {
    for (int i = 0; i < nsteps; i++) ((Run *)h)->s->Advance(dt);
}

void vxo_advance_fields(void *h, double dt, int nsteps) {
    for (int i = 0; i < nsteps; i++) ((Run *)h)->s->AdvanceFields(dt);
}

// One RK stage of SolverManager::Advance (SolverManager.cpp:29-38), and the
// pieces of it, for stage-by-stage parity checks.
void vxo_stage_moments(void *h) { ((Run *)h)->s->em->AssembleRhoAndJ(); }
void vxo_stage_poisson(void *h) { ((Run *)h)->s->em->UpdatePotential(); }
void vxo_stage_vlasov(void *h, double dt, int step) {
    for (auto &m : ((Run *)h)->s->meshes) m->Advance(dt, step);
}
void vxo_stage_vlasov_sub(void *h, double dt, int step, int sub) {
    Run *r = (Run *)h;
    for (auto &m : r->s->meshes) {
        for (int i = (int)m->levels.size() - 1; i > -1; i--) m->levels[i]->FCTTimeStep(dt, step, sub);
        if (sub == 0) m->PushData(2);
        if (sub == 1) m->PushBoundaryC();
        if (sub == 2) m->PushData(1);
    }
}
void vxo_stage_fields(void *h, double dt, int step) {
    Run *r = (Run *)h;
    r->s->settings.UpdateTime(step, dt);
    r->s->em->RGKStep(step, dt);
}

// Named arrays.  Patch planes (base patch of species 0) are returned with
// `ghosts` ghost layers (0..2), row-major over (x, p).
// Plane `what` of patch p, or -1 when the name is not a plane.
static int patch_plane(orc::Patch &p, const std::string &w, int ghosts, double *out, int n) {
    if (w == "f0") return copy_plane(p, p.f[0], ghosts, out, n);
    if (w == "f1") return copy_plane(p, p.f[1], ghosts, out, n);
    if (w == "f2") return copy_plane(p, p.f[2], ghosts, out, n);
    if (w == "FxDS") return copy_plane(p, p.FxDS, ghosts, out, n);
    if (w == "FpDS") return copy_plane(p, p.FpDS, ghosts, out, n);
    if (w == "Rp") return copy_plane(p, p.Rp, ghosts, out, n);
    if (w == "Rm") return copy_plane(p, p.Rm, ghosts, out, n);
    if (w == "Cx") return copy_plane(p, p.Cx, ghosts, out, n);
    if (w == "Cp") return copy_plane(p, p.Cp, ghosts, out, n);
    for (int k = 0; k < 6; k++) {
        if (w == "FxH" + std::to_string(k)) return copy_plane(p, p.FxH[k], ghosts, out, n);
        if (w == "FpH" + std::to_string(k)) return copy_plane(p, p.FpH[k], ghosts, out, n);
        if (w == "FxL" + std::to_string(k)) return copy_plane(p, p.FxL[k], ghosts, out, n);
        if (w == "FpL" + std::to_string(k)) return copy_plane(p, p.FpL[k], ghosts, out, n);
    }
    if (w == "nested") {
        std::vector<double> v(p.is_nested.begin(), p.is_nested.end());
        return copy_plane(p, v, ghosts, out, n);
    }
    return -1;
}

int vxo_get(void *h, const char *what, int ghosts, double *out, int n) {
    Run *r = (Run *)h;
    orc::Field &em = *r->s->em;
    orc::Patch &p = base(r);
    std::string w(what);
    int m = patch_plane(p, w, ghosts, out, n);
    if (m >= 0) return m;
    if (w == "phi") return copy_out(em.PHI, out, n);
    if (w == "charge") return copy_out(em.charge, out, n);
    // charge of species i from the last assembly (EMSolver.cpp:104-122)
    if (w.rfind("charge_sp", 0) == 0) {
        unsigned sp = (unsigned)std::stoi(w.substr(9));
        return sp < em.charges.size() ? copy_out(em.charges[sp], out, n) : -1;
    }
    if (w == "J") return copy_out(em.J, out, n);
    if (w == "neutral") return copy_out(em.neutralizationCharge, out, n);
    if (w == "a2") return copy_out(em.a_squared, out, n);
    if (w == "By") return copy_out(em.By, out, n);
    if (w == "Bz") return copy_out(em.Bz, out, n);
    if (w == "Ey") return copy_out(em.Ey, out, n);
    if (w == "Ez") return copy_out(em.Ez, out, n);
    if (w == "Ay") return copy_out(em.Ay, out, n);
    if (w == "Az") return copy_out(em.Az, out, n);
    if (w == "E") {
        std::vector<double> e(em.x_size);
        for (unsigned i = 0; i < em.x_size; i++) e[i] = em.GetEfield(i);
        return copy_out(e, out, n);
    }
    if (w == "dNdP") {
        em.AssembleEnergy();
        return copy_out(em.energies[0], out, n);
    }
    if (w == "scalars") {  // Ex0, time, dx, dp, pmin, mass, charge, n_prepad
        std::vector<double> s = {em.Ex0, r->s->settings.time, p.dx, p.dp, r->s->settings.pmin[0],
                                 r->s->settings.m[0], r->s->settings.q[0], (double)em.n_prepad};
        return copy_out(s, out, n);
    }
    return -1;
}

// Patch layout of species sp: 9 ints per patch (level with 0 the finest,
// x_pos, p_pos, n_x, n_p, up, down, left, right), levels finest first, patches
// in the reference's order.  Returns the number of ints.
int vxo_layout(void *h, int sp, int *out, int n) {
    Run *r = (Run *)h;
    std::vector<int> v;
    auto &levels = r->s->meshes.at(sp)->levels;
    for (unsigned l = 0; l < levels.size(); l++)
        for (auto &p : levels[l]->rectangles) {
            int e[9] = {(int)l, p->x_pos, p->p_pos, p->n_x, p->n_p, p->up, p->down, p->left, p->right};
            v.insert(v.end(), e, e + 9);
        }
    if (out && n >= (int)v.size()) std::memcpy(out, v.data(), sizeof(int) * v.size());
    return (int)v.size();
}

// Plane of patch `idx` of level `lvl` (0 = finest) of species sp.
int vxo_patch_get(void *h, int sp, int lvl, int idx, const char *what, int ghosts, double *out, int n) {
    Run *r = (Run *)h;
    auto &levels = r->s->meshes.at(sp)->levels;
    if (lvl < 0 || lvl >= (int)levels.size() || idx < 0 || idx >= (int)levels[lvl]->rectangles.size())
        return -2;
    return patch_plane(*levels[lvl]->rectangles[idx], what, ghosts, out, n);
}

// Flags of hierarchy level lvl (0 = coarsest) from the current data
// (Mesh::getError, Mesh.cpp:212-295), as (i, j) pairs.  Returns 2 * count.
int vxo_flags(void *h, int sp, int lvl, int *out, int n) {
    Run *r = (Run *)h;
    std::vector<orc::coords> fl;
    r->s->meshes.at(sp)->getError(lvl, false, fl);
    if (out && n >= 2 * (int)fl.size())
        for (size_t k = 0; k < fl.size(); k++) {
            out[2 * k] = fl[k].first;
            out[2 * k + 1] = fl[k].second;
        }
    return 2 * (int)fl.size();
}

// SolverManager::reGrid (SolverManager.cpp:48-58)
void vxo_regrid(void *h) { ((Run *)h)->s->reGrid(); }

// For the pins of the Vlasov advance by exact solutions: set slots 0 and 1
// of the base patch's interior (n_x x n_p, row-major) of species 0; ghosts
// are left as they are (zero at the walls).  Returns -1 on a size mismatch.
int vxo_set_f(void *h, const double *f, int n) {
    orc::Patch &p = base((Run *)h);
    if (n != p.n_x * p.n_p) return -1;
    for (int i = 0; i < p.n_x; i++)
        for (int j = 0; j < p.n_p; j++) {
            p.f[0][p.NS(i, j)] = f[i * p.n_p + j];
            p.f[1][p.NS(i, j)] = f[i * p.n_p + j];
        }
    return 0;
}

// The Vlasov part of SolverManager::Advance (SolverManager.cpp:28-39) with an
// imposed field: at stage i the stage clock is advanced first (Settings.cpp:
// 166-179, so it reads t_n + c_i dt), then the potential is zeroed and Ex0 set
// to E0 cos(omega t), so that every cell sees that uniform field
// (EMSolver.cpp:137-154).  The transverse fields stay as they are (zero unless
// a drive ran); the moments, the Poisson solve and the Maxwell stage are
// skipped.
void vxo_vlasov_imposed(void *h, double dt, int nsteps, double E0, double omega) {
    Run *r = (Run *)h;
    orc::Field &em = *r->s->em;
    for (int n = 0; n < nsteps; n++)
        for (int i = 0; i < 6; i++) {
            r->s->settings.UpdateTime(i, dt);
            std::fill(em.PHI.begin(), em.PHI.end(), 0.0);
            em.Ex0 = E0 * std::cos(omega * r->s->settings.time);
            for (auto &m : r->s->meshes) m->Advance(dt, i);
        }
}

// Mesh::splitRectangle (Mesh.cpp:644-792) on box (x0, p0)-(x1, p1) with n
// flagged cells (x, p pairs); writes 4 ints per resulting box, returns 4 * count.
int vxo_split(int x0, int p0, int x1, int p1, const int *flags, int n, double minEfficiency, int *out,
              int nout) {
    std::vector<orc::coords> fl;
    for (int k = 0; k < n; k++) fl.push_back({flags[2 * k], flags[2 * k + 1]});
    orc::level boxes = orc::Mesh::splitRectangle({{x0, p0}, {x1, p1}}, fl, minEfficiency);
    if (out && nout >= 4 * (int)boxes.size())
        for (size_t b = 0; b < boxes.size(); b++) {
            out[4 * b] = boxes[b].first.first;
            out[4 * b + 1] = boxes[b].first.second;
            out[4 * b + 2] = boxes[b].second.first;
            out[4 * b + 3] = boxes[b].second.second;
        }
    return 4 * (int)boxes.size();
}

// The sub-cell interpolants of the base patch of species 0 from five
// neighbouring cell averages f[0..4] (cells -2..2): which = 0 for
// GetInterpolantsREF (r sub-cells, Rectangle.cpp:121-137), 1 for
// GetInterpolantsREL (relativeToBottom sub-cells, Rectangle.cpp:139-155).
// Returns the number of values.
int vxo_interp(void *h, int which, const double *f, double *out, int n) {
    orc::Patch &p = base((Run *)h);
    std::vector<double> t = which == 0 ? p.GetInterpolantsREF(f[0], f[1], f[2], f[3], f[4])
                                       : p.GetInterpolantsREL(f[0], f[1], f[2], f[3], f[4]);
    return copy_out(t, out, n);
}

// Standalone pieces for unit tests.
double vxo_weno_nomax(double f1, double f2, double f3, double f4, int right) {
    return orc::weno_nomax(f1, f2, f3, f4, right != 0);
}
double vxo_face_weno(double f1, double f2, double f3, double f4) {
    return orc::face_weno(f1, f2, f3, f4);
}
void vxo_poisson(void *h, const double *b, double *x) {
    Run *r = (Run *)h;
    std::vector<double> bv(b, b + r->s->em->x_size), xv;
    r->s->em->PoissonSolve(bv, xv);
    std::memcpy(x, xv.data(), sizeof(double) * xv.size());
}
void vxo_poisson_approx(void *h, const double *b, double *x) {
    Run *r = (Run *)h;
    std::vector<double> bv(b, b + r->s->em->x_size), xv;
    r->s->em->PoissonApprox(bv, xv);
    std::memcpy(x, xv.data(), sizeof(double) * xv.size());
}
double vxo_log(double x) { return vx_log(x); }
double vxo_k1e(double z) { return vx::Scenario::k1e(z); }
}
