// Host classes of the MI355X framework (see veritas.hpp).  Every numerical
// step is a call into the HIP library through the C ABI; this file owns
// orchestration, device memory and the user hooks, as the reference's classes
// do on the CPU.
#include "veritas.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iomanip>
#include <sstream>

namespace vxh {

// ---------------------------------------------------------------- Settings
Hooks Hooks::FromC(const vx_hooks &h) {
    Hooks k;
    void *u = h.user;
    if (h.initial_distribution) {
        auto f = h.initial_distribution;
        k.InitialDistribution = [f, u](double x, double p, int sp) { return f(u, x, p, sp); };
    }
    if (h.get_by) {
        auto f = h.get_by;
        k.GetBY = [f, u](double x, double t) { return f(u, x, t); };
    }
    if (h.get_bz) {
        auto f = h.get_bz;
        k.GetBZ = [f, u](double x, double t) { return f(u, x, t); };
    }
    if (h.refinement_override) {
        auto f = h.refinement_override;
        k.RefinementOverride = [f, u](double x, double p, int d, int sp) { return f(u, x, p, d, sp) != 0; };
    }
    if (h.settings_override) {
        auto f = h.settings_override;
        k.settingsOverride = [f, u](Settings &s) {
            vx_settings_desc d = s.Describe();
            f(u, &d);
            // the values the C description carries go back into the settings
            s.minEfficiency = d.minEfficiency;
            s.dx = d.dx;
            s.refinementCriteria = d.refinementCriteria;
            s.cfl = d.cfl;
            s.sizeWeight = d.sizeWeight;
            s.preLength = d.preLength;
            s.postLength = d.postLength;
            s.x_size = d.nx;
            s.refinementRatio = d.r;
            s.maxDepth = (int)d.Lfinest - 1;
            s.quadratureDepth = d.quadratureDepth;
            if (d.n_species < 1 || d.n_species > VX_MAX_SPECIES)
                throw Error("settingsOverride: n_species out of range");
            s.m.assign(d.mass, d.mass + d.n_species);
            s.q.assign(d.charge, d.charge + d.n_species);
            s.dp.assign(d.dp, d.dp + d.n_species);
            s.pmin.assign(d.pmin, d.pmin + d.n_species);
            s.p_size.assign(d.np, d.np + d.n_species);
        };
    }
    return k;
}

static Input input_of(const vx_settings_desc &d) {
    Input g;
    g.minEfficiency = d.minEfficiency;
    g.dx = d.dx;
    g.refinementCriteria = d.refinementCriteria;
    g.cfl = d.cfl;
    g.sizeWeight = d.sizeWeight;
    g.preLength = d.preLength;
    g.postLength = d.postLength;
    g.nx = d.nx;
    g.r = d.r;
    g.Lfinest = d.Lfinest;
    g.quadratureDepth = d.quadratureDepth;
    return g;
}

static Particles particles_of(const vx_settings_desc &d) {
    if (d.n_species < 1 || d.n_species > VX_MAX_SPECIES) throw Error("n_species must be 1..8");
    Particles pt;
    pt.mass.assign(d.mass, d.mass + d.n_species);
    pt.charge.assign(d.charge, d.charge + d.n_species);
    pt.dp.assign(d.dp, d.dp + d.n_species);
    pt.pmin.assign(d.pmin, d.pmin + d.n_species);
    pt.np.assign(d.np, d.np + d.n_species);
    pt.misc.assign(d.n_species, {});
    return pt;
}

// Settings.cpp:5-31: copy the inputs, then let the user override them
Settings::Settings(const Input &grid, const Particles &particles, const Output &out, Hooks h)
    : output(out), hooks(std::move(h)) {
    if (!hooks.InitialDistribution) throw Error("Settings: the InitialDistribution hook is required");
    dx = grid.dx;
    dp = particles.dp;
    maxDepth = (int)grid.Lfinest - 1;
    x_size = grid.nx;
    p_size = particles.np;
    refinementRatio = grid.r;
    refinementCriteria = grid.refinementCriteria;
    minEfficiency = grid.minEfficiency;
    preLength = grid.preLength;
    postLength = grid.postLength;
    cfl = grid.cfl;
    sizeWeight = grid.sizeWeight;
    m = particles.mass;
    q = particles.charge;
    pmin = particles.pmin;
    temp = particles.misc;
    tempEM = grid.tempEM;
    time = 0.0;
    quadratureDepth = grid.quadratureDepth;
    if (hooks.settingsOverride) hooks.settingsOverride(*this);
    finish();
}

Settings::Settings(const vx_settings_desc &d, const vx_hooks &h)
    : Settings(input_of(d), particles_of(d), Output(), Hooks::FromC(h)) {}

Settings::Settings(const vx::Scenario &s) : Settings(s.desc(), s.hooks()) {
    // the hooks point at the scenario: keep a copy alive and re-point them
    auto own = std::make_shared<const vx::Scenario>(s);
    hooks = Hooks::FromC(own->hooks());
    scenario_ = own;
    identity = {(double)s.kind, s.n_ref, s.length, s.pmin, s.pmax, s.theta, s.amplitude, s.k, s.drift,
                (double)s.modes, (double)s.seed, s.a0, s.n_over_nc, s.ramp, s.fwhm_cycles, s.lambda0, s.gap};
}

vx_settings_desc Settings::Describe() const {
    vx_settings_desc d{};
    d.minEfficiency = minEfficiency;
    d.dx = dx;
    d.refinementCriteria = refinementCriteria;
    d.cfl = cfl;
    d.sizeWeight = sizeWeight;
    d.preLength = preLength;
    d.postLength = postLength;
    d.nx = x_size;
    d.r = refinementRatio;
    d.Lfinest = (uint32_t)(maxDepth + 1);
    d.quadratureDepth = quadratureDepth;
    d.n_species = (int32_t)std::min<size_t>(q.size(), VX_MAX_SPECIES);
    for (int i = 0; i < d.n_species; i++) {
        d.mass[i] = m[i];
        d.charge[i] = q[i];
        d.dp[i] = dp[i];
        d.pmin[i] = pmin[i];
        d.np[i] = p_size[i];
    }
    return d;
}

// Settings.cpp:35-58
void Settings::finish() {
    time = 0.0;
    if (refinementRatio % 2) throw Error("Mesh refinement ratio 'refinementRatio' must be an even value.");
    if (x_size % refinementRatio != 0)
        throw Error("'x_size' must be completely divisible by the Mesh refinement ratio 'refinementRatio'.");
    for (unsigned p = 0; p < p_size.size(); p++)
        if (p_size[p] % refinementRatio != 0)
            throw Error("'p_size' at index (" + std::to_string(p) +
                        ") must be completely divisible by the Mesh refinement ratio 'refinementRatio'.");
    if (m.size() != q.size() || dp.size() != q.size() || pmin.size() != q.size() || p_size.size() != q.size())
        throw Error("Particles: mass, charge, np, dp and pmin must have one entry per species");
    if (q.empty() || q.size() > VX_MAX_SPECIES) throw Error("Particles: 1 to 8 species");
    if (temp.size() < q.size()) temp.resize(q.size());
    x_size_finest = x_size * std::pow(refinementRatio, maxDepth);
    p_size_finest.clear();
    for (unsigned i = 0; i < p_size.size(); i++)
        p_size_finest.push_back(p_size[i] * std::pow(refinementRatio, maxDepth));
    fMax.assign(q.size(), 0.0);
    DetermineMaximum();
}

// Settings.cpp:60-125 (OUTW = 80, veritas.hpp:30)
std::string Settings::Banner() const {
    const int W = 80;
    std::ostringstream o;
    auto title = [&](const std::string &t, char sp) {
        int space = W - (int)t.size();
        int l = space / 2, r = space - l;
        o << std::string(std::max(l, 0), sp) << t << std::string(std::max(r, 0), sp) << "\n";
    };
    int maxleft = 29, maxright = 25;
    bool nonneg = false;
    for (double v : pmin) nonneg = nonneg || (int)v >= 0;
    if (nonneg) maxleft--;
    title("VERITAS. Current build SHA1: veritas_amd (MI355X)", ' ');
    o << "\n";
    title(" Configuration Parameters ", '=');
    auto row = [&](const std::string &a, const std::string &av, const std::string &b, const std::string &bv) {
        o << std::left << std::setw(maxleft) << a << std::setw(W / 2 - maxleft) << av;
        if (!b.empty()) o << std::setw(maxright) << b << std::setw(W / 2 - maxright) << bv;
        o << "\n";
    };
    auto num = [](double v) {
        std::ostringstream t;
        t << std::setprecision(5) << v;
        return t.str();
    };
    row("Minimum Efficiency: ", num(minEfficiency), "Mesh refinement ratio: ", num(refinementRatio));
    row("Maximum number of levels: ", num(maxDepth + 1), "Error tolerance: ", num(refinementCriteria));
    row("nx: ", num(x_size), "PreLength: ", num(preLength));
    row("dx: ", num(dx), "PostLength: ", num(postLength));
    title(" Particle Properties ", '=');
    for (size_t i = 0; i < q.size(); i++) {
        row("np: ", num(p_size[i]), "Mass: ", num(m[i]));
        row("dp: ", num(dp[i]), "Charge: ", num(q[i]));
        o << std::left << std::setw(pmin[i] >= 0 ? maxleft : maxleft - 1) << "Momentum Cutoff: "
          << std::setw(W / 2 - maxleft) << num(pmin[i]) << "\n";
        o << std::setw(maxleft) << "Misc values: ";
        for (size_t j = 0; j < temp[i].size(); j++) o << num(temp[i][j]) << (j + 1 < temp[i].size() ? "; " : "");
        o << "\n";
        if (q.size() > 1 && i + 1 < q.size()) o << std::string(80, '-') << "\n";
    }
    title(" Output Settings ", '=');
    auto b = [](bool v) { return std::string(v ? "true" : "false"); };
    row("Rectangle Data: ", b(output.rectangleData), "Charge: ", b(output.charge));
    row("Longitudinal ElectricField: ", b(output.EFieldLongitudinal), "Energy: ", b(output.energy));
    row("Transverse ElectricField: ", b(output.EFieldTransverse), "Electrostatic Potential: ", b(output.potential));
    row("Transverse MagniticField: ", b(output.BFieldTransverse), "Squared Vector Potental: ", b(output.AFieldSquared));
    row("Output Precision: ", num(output.precision), "", "");
    o << std::string(80, '=') << "\n\n";
    return o.str();
}

double Settings::GetDp(int level, int sp) const {
    return std::pow(refinementRatio, level - maxDepth) * dp.at(sp);
}
double Settings::GetDx(int level) const { return std::pow(refinementRatio, level - maxDepth) * dx; }
int Settings::GetXSize(int level) const { return x_size * std::pow(refinementRatio, -level + maxDepth); }
int Settings::GetPSize(int level, int sp) const {
    return p_size[sp] * std::pow(refinementRatio, -level + maxDepth);
}

// Settings.cpp:166-179
void Settings::UpdateTime(int step, double dt) {
    if (step == 1) time += (0.5 * dt);
    else if (step == 2) time += (0.332 - 0.5) * dt;
    else if (step == 3) time += (0.62 - 0.332) * dt;
    else if (step == 4) time += (0.85 - 0.62) * dt;
    else if (step == 5) time += (1.0 - 0.85) * dt;
}

// Settings.cpp:181-195
void Settings::DetermineMaximum() {
    double d = GetDx(0);
    double mValue(0.0);
    for (unsigned j = 0; j < q.size(); j++) {
        for (int i = 0; i < GetXSize(0); i++) {
            double t(InitialDistribution((i + 0.5) * d, 0, j));
            mValue = t > mValue ? t : mValue;
        }
        fMax[j] = mValue;
    }
}

// ---------------------------------------------------------------- Device
Device::Device(int device) {
    int rc = vx_ctx_create(device, &ctx);
    if (rc) throw Error("vx_ctx_create failed (no usable GPU?)");
}
Device::~Device() { vx_ctx_destroy(ctx); }
void Device::check(int rc, const char *what) const {
    if (rc) throw Error(std::string(what) + ": " + vx_last_error(ctx));
}

// ---------------------------------------------------------------- Rectangle
Rectangle::Rectangle(int n_x_, int n_p_, int x_pos_, int p_pos_, int depth_, Settings &s,
                     std::shared_ptr<Device> d, bool up_, bool down_, bool left_, bool right_,
                     int sp)
    : n_x(n_x_), n_p(n_p_), x_pos(x_pos_), p_pos(p_pos_), depth(depth_), particleType(sp),
      up(up_), down(down_), left(left_), right(right_), settings(s), dev(std::move(d)) {
    relativeToBottom = std::pow(s.refinementRatio, depth);
    P.n_x = n_x;
    P.n_p = n_p;
    P.x_pos = x_pos;
    P.p_pos = p_pos;
    P.rtb = (int)relativeToBottom;
    P.pitch = vx_plane_pitch(n_p);
    P.left = left;
    P.right = right;
    P.down = down;
    P.up = up;
    P.dx = s.GetDx(depth);
    P.dp = s.GetDp(depth, sp);
    P.pmin = s.pmin.at(sp);
    P.q = s.GetCharge(sp);
    P.m = s.GetMass(sp);
    // as Rectangle.cpp:1271-1275 and Rectangle.hpp:182 evaluate them
    P.dx_inv = 1 / P.dx;
    P.dp_inv = 1 / P.dp;
    P.cc = vx_cs * vx_cs * P.m;
    double m_inv = 1 / P.m;
    P.mc2inv = (m_inv * vx_c_inv) * (m_inv * vx_c_inv);
    for (int k = 0; k < 3; k++) P.f[k] = buf[k] = alloc_plane();
    for (int k = 0; k < 6; k++) {
        P.FxH[k] = alloc_plane();
        P.FpH[k] = alloc_plane();
    }
    P.FxL0 = alloc_plane();
    P.FpL0 = alloc_plane();
    P.FxDS = alloc_plane();
    P.FpDS = alloc_plane();
    P.Rp = alloc_plane();
    P.Rm = alloc_plane();
    P.Cx = alloc_plane();
    P.Cp = alloc_plane();
    P.nested = nullptr;
    // adaptive-mesh links: sentinel neighbours, no finer patches
    // (Rectangle.cpp:40-66); error weights (Rectangle.cpp:68-74)
    const int r = s.refinementRatio;
    const size_t NSsz = (size_t)(n_x + 4) * (n_p + 4);
    ext_xm.assign(n_p / r + 2, nullptr);
    ext_xp.assign(n_p / r + 2, nullptr);
    ext_pm.assign(n_x / r, nullptr);
    ext_pp.assign(n_x / r, nullptr);
    same_xm.assign(n_p / r + 2, 1);
    same_xp.assign(n_p / r + 2, 1);
    same_pm.assign(n_x / r, 1);
    same_pp.assign(n_x / r, 1);
    finer_level.assign(NSsz, nullptr);
    finer_level_x.assign(NSsz, nullptr);
    finer_level_p.assign(NSsz, nullptr);
    is_nested.assign(NSsz, 0);
    is_ilb_x.assign(NSsz, 0);
    is_ilb_p.assign(NSsz, 0);
    is_interpolated.assign(NSsz, 0);
    const double fmx = s.GetfMax(sp);
    errorWeights[0] = 0.5 / fmx * std::pow(r, -0.5 * depth);
    errorWeights[1] = 0.5 / fmx * std::pow(r, -0.5 * depth);
    errorWeights[2] = 0.5 / fmx * std::pow(r, -2.0 * depth);
    errorWeights[3] = 0.5 / fmx * std::pow(r, -2.0 * depth);
    errorWeights[4] = s.sizeWeight / fmx * std::pow(r, -s.maxDepth + depth);
    void *c = nullptr;
    dev->check(vx_malloc(dev->ctx, sizeof(double) * n_x * P.rtb, &c), "alloc chargeR");
    allocations.push_back(c);
    P.chargeR = (double *)c;
    dev->check(vx_malloc(dev->ctx, sizeof(double) * n_x * P.rtb, &c), "alloc currentR");
    allocations.push_back(c);
    P.currentR = (double *)c;
}

Rectangle::~Rectangle() {
    for (void *p : allocations) vx_free(dev->ctx, p);
}

double *Rectangle::alloc_plane() {
    void *p = nullptr;
    dev->check(vx_malloc(dev->ctx, sizeof(double) * vx_plane_elems(n_x, n_p), &p), "alloc plane");
    allocations.push_back(p);
    return (double *)p;
}

// Rectangle.cpp:616-665: nvals^2 midpoint quadrature of the user hook, on
// the host as in the reference; the cell averages are then uploaded.
void Rectangle::InitializeDistribution() {
    const int nvals = std::pow(settings.refinementRatio, depth + settings.quadratureDepth);
    const int pitch = P.pitch;
    std::vector<double> h(vx_plane_elems(n_x, n_p), 0.0);
    const double dxl = P.dx, dpl = P.dp, pmin = P.pmin;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n_x; i++) {
        for (int j = 0; j < n_p; j++) {
            double temp = 0.0, xp = 0.0, pp = 0.0;
            for (int k = 0; k < nvals; k++) {
                for (int l = 0; l < nvals; l++) {
                    xp = ((0.5 + k) / nvals + i + x_pos) * dxl;
                    pp = pmin + dpl * (((0.5 + l) / nvals + j) + p_pos);
                    temp += settings.InitialDistribution(xp, pp, particleType);
                }
            }
            h[(size_t)(i + 2) * pitch + (j + 2)] = temp * (1 / std::pow(nvals, 2.0));
        }
    }
    size_t bytes = sizeof(double) * h.size();
    for (double v : h) fScale = std::max(fScale, std::fabs(v));
    dev->check(vx_memcpy_h2d(dev->ctx, P.f[0], h.data(), bytes), "upload f0");
    dev->check(vx_memcpy_h2d(dev->ctx, P.f[1], h.data(), bytes), "upload f1");
}

void Rectangle::FCTTimeStep(double timestep, int step, int subStep) {
    vx_ctx *c = dev->ctx;
    UnaliasSlots();
    switch (subStep) {
        case 0: dev->check(vx_fct_sub0(c, &P, step, timestep, &EMSolver->fields()), "fct sub0"); break;
        case 1: dev->check(vx_fct_sub1(c, &P), "fct sub1"); break;
        case 2: dev->check(vx_fct_sub2(c, &P), "fct sub2"); break;
        case 3: dev->check(vx_fct_sub3(c, &P), "fct sub3"); break;
        default: throw Error("FCTTimeStep: bad sub-step");
    }
}

void Rectangle::FCTStageWalls(double timestep, int step) {
    UnaliasSlots();
    dev->check(vx_fct_stage_walls(dev->ctx, &P, step, timestep, &EMSolver->fields()), "fct stage");
}

// f[1] = f_s is read, f_{s+1} goes to the plane that holds neither f^n nor
// f_s; at stage 5 f^n and f_s become the same plane (f[0] <- f[1],
// Rectangle.cpp:1614-1622, without a copy).
void Rectangle::FCTStageFused(double timestep, int step, bool mixed) {
    double *out = nullptr;
    for (double *b : buf)
        if (b != P.f[0] && b != P.f[1]) {
            out = b;
            break;
        }
    if (mixed)
        dev->check(vx_fct_stage_fused_mp(dev->ctx, &P, step, timestep, &EMSolver->fields(), P.f[1], out,
                                         fScale > 0.0 ? fScale : settings.GetfMax(particleType)),
                   "fct stage fused (mixed)");
    else
        dev->check(vx_fct_stage_fused(dev->ctx, &P, step, timestep, &EMSolver->fields(), P.f[1], out),
                   "fct stage fused");
    P.f[1] = out;
    if (step == 5) P.f[0] = out;
    for (double *b : buf)
        if (b != P.f[0] && b != P.f[1]) P.f[2] = b;
    slot_version++;
}

void Rectangle::FusedItem(double timestep, int step, bool mixed, vx_stage_item *item) {
    double *out = nullptr;
    for (double *b : buf)
        if (b != P.f[0] && b != P.f[1]) {
            out = b;
            break;
        }
    item->p = P;
    item->fl = EMSolver->fields();
    item->fin = P.f[1];
    item->fout = out;
    item->dt = timestep;
    item->f_scale = mixed ? (fScale > 0.0 ? fScale : settings.GetfMax(particleType)) : 0.0;
    P.f[1] = out;
    if (step == 5) P.f[0] = out;
    for (double *b : buf)
        if (b != P.f[0] && b != P.f[1]) P.f[2] = b;
    slot_version++;
}

void Rectangle::UnaliasSlots() {
    if (P.f[0] != P.f[1]) return;
    double *spare[2];
    int n = 0;
    for (double *b : buf)
        if (b != P.f[0]) spare[n++] = b;
    P.f[1] = spare[0];
    P.f[2] = spare[1];
    slot_version++;
    dev->check(vx_memcpy_d2d(dev->ctx, P.f[1], P.f[0], sizeof(double) * vx_plane_elems(n_x, n_p)),
               "unalias slots");
}

void Rectangle::CalculateRhoAndJ() {
    if (EMSolver->compensated_moments)
        dev->check(vx_moments_patch_compensated(dev->ctx, &P, &EMSolver->fields()), "moments");
    else
        dev->check(vx_moments_patch(dev->ctx, &P, &EMSolver->fields()), "moments");
}

std::vector<double> Rectangle::DownloadSlot(int slot, int g) const {
    std::vector<double> out((size_t)(n_x + 2 * g) * (n_p + 2 * g));
    dev->check(vx_patch_download(dev->ctx, &P, slot, g, out.data()), "download slot");
    return out;
}

std::vector<double> Rectangle::Download(const double *plane, int g) const {
    std::vector<double> full(vx_plane_elems(n_x, n_p));
    dev->check(vx_memcpy_d2h(dev->ctx, full.data(), plane, sizeof(double) * full.size()), "download");
    int nx = n_x + 2 * g, np = n_p + 2 * g;
    std::vector<double> out((size_t)nx * np);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < np; j++)
            out[(size_t)i * np + j] = full[(size_t)(i - g + 2) * P.pitch + (j - g + 2)];
    return out;
}

const double *Rectangle::Plane(const std::string &w) const {
    if (w == "f0") return P.f[0];
    if (w == "f1") return P.f[1];
    if (w == "f2") return P.f[2];
    if (w == "FxDS") return P.FxDS;
    if (w == "FpDS") return P.FpDS;
    if (w == "Rp") return P.Rp;
    if (w == "Rm") return P.Rm;
    if (w == "Cx") return P.Cx;
    if (w == "Cp") return P.Cp;
    if (w == "FxL0") return P.FxL0;
    if (w == "FpL0") return P.FpL0;
    for (int k = 0; k < 6; k++) {
        if (w == "FxH" + std::to_string(k)) return P.FxH[k];
        if (w == "FpH" + std::to_string(k)) return P.FpH[k];
    }
    return nullptr;
}

// ---------------------------------------------------------------- Level
Level::Level(int sp, int depth_, Settings &s) : depth(depth_), particleType(sp), settings(s) {}

void Level::FCTTimeStep(double timestep, int step, int subStep) {
    for (auto &r : rectangles) r->FCTTimeStep(timestep, step, subStep);
}

// ---------------------------------------------------------------- Mesh
void Mesh::SetFieldSolver(EMFieldSolver *solver) {
    EMSolver = solver;
    for (auto &l : levels)
        for (auto &r : l->rectangles) r->SetFieldSolver(solver);
}

void Mesh::Sub0(double timeStep, int step) {
    for (auto &l : levels)
        for (auto &r : l->rectangles) r->UnaliasSlots();
    EnsureTable();
    const vx_fields &fl = EMSolver->fields();
    // every patch's fluxes in one launch (patches write only their own planes)
    if (mixed)
        dev->check(vx_fct_sub0_fluxes_table_mp(dev->ctx, table, step, timeStep, &fl, FScale()),
                   "fct sub0 fluxes (mixed)");
    else
        dev->check(vx_fct_sub0_fluxes_table(dev->ctx, table, step, timeStep, &fl), "fct sub0 fluxes");
    for (unsigned i = 1; i < levels.size(); i++) levels[i]->RunFluxProgram(step, timeStep, fl);
    dev->check(vx_fct_sub0_predictor_table(dev->ctx, table, step, timeStep, fl.n_f), "fct sub0 predictor");
}

double Mesh::FScale() const {
    double s = 0.0;
    for (const auto &l : levels)
        for (const auto &r : l->rectangles)
            s = std::max(s, r->fScale > 0.0 ? r->fScale : settings.GetfMax(particleType));
    return s;
}

void Mesh::EnsureTable() {
    int v = 0;
    for (Rectangle *r : order) v += r->slot_version;
    if (v != table_version) {
        RefreshTable();
        table_version = v;
    }
}

// Mesh.cpp:64-89.  Sub-step 0 is split around the coarse-fine flux programs:
// every patch's fluxes, then the coarse faces on finer patches' edges (finest
// coarse level first), then the low-order predictors.
void Mesh::Advance(double timeStep, int step) {
    Rectangle *single = SinglePatch() ? levels.back()->rectangles[0].get() : nullptr;
    if (fused && single && single->AllWalls()) {
        single->FCTStageFused(timeStep, step, mixed);
        return;
    }
    const int n_f = EMSolver->fields().n_f;
    if (!batched) {
        // the plain per-patch form: FP64 sub-step launches one patch at a time,
        // the same task programs between them
        const vx_fields &fl = EMSolver->fields();
        auto each = [&](const std::function<int(const vx_patch *)> &f, const char *what) {
            for (auto &l : levels)
                for (auto &r : l->rectangles) dev->check(f(&r->desc()), what);
        };
        for (auto &l : levels)
            for (auto &r : l->rectangles) r->UnaliasSlots();
        each([&](const vx_patch *p) { return vx_fct_sub0_fluxes(dev->ctx, p, step, timeStep, &fl); }, "sub0 fluxes");
        for (unsigned i = 1; i < levels.size(); i++) levels[i]->RunFluxProgram(step, timeStep, fl);
        each([&](const vx_patch *p) { return vx_fct_sub0_predictor(dev->ctx, p, step, timeStep); }, "predictor");
        PushData(2);
        each([&](const vx_patch *p) { return vx_fct_sub1(dev->ctx, p); }, "sub1");
        PushBoundaryC();
        each([&](const vx_patch *p) { return vx_fct_sub2(dev->ctx, p); }, "sub2");
        PushData(1);
        if (step == 5) each([&](const vx_patch *p) { return vx_fct_sub3(dev->ctx, p); }, "sub3");
        return;
    }
    // each sub-step is one launch over the mesh's patch table (Level::FCTTimeStep
    // over every level, Mesh.cpp:64-89; the patches of a sub-step are independent)
    Sub0(timeStep, step);
    PushData(2);
    EnsureTable();
    // sub-steps 1 and 2 in one launch, then the limiter-C synchronisation
    // with the cells next to the faces it changes corrected again
    if (mixed) dev->check(vx_fct_sub12_table_mp(dev->ctx, table, n_f, FScale()), "fct sub1+2 (mixed)");
    else dev->check(vx_fct_sub12_table(dev->ctx, table, n_f), "fct sub1+2");
    if (pushcc_prog)
        dev->check(vx_program_run(dev->ctx, table, pushcc_prog, VX_MODE_GHOST, 1, settings.refinementRatio),
                   "limiter sync");
    PushData(1);
    if (step == 5) {
        EnsureTable();
        dev->check(vx_fct_sub3_table(dev->ctx, table, n_f), "fct sub3");
    }
}

// Mesh.cpp:91-106: level 0's type 1, then types 0 and 1 of the coarser levels,
// the coarsest level's type 3, and type 2 from the second-coarsest level to
// the finest, run as one merged program (built in BuildPrograms)
void Mesh::PushData(int val) {
    EnsureTable();
    if (push_prog)
        dev->check(vx_program_run(dev->ctx, table, push_prog, VX_MODE_GHOST, val, settings.refinementRatio),
                   "push data");
}

// Mesh.cpp:904-917: types 4, 5 and 6 over every level, one merged program
void Mesh::PushBoundaryC() {
    EnsureTable();
    if (pushc_prog)
        dev->check(vx_program_run(dev->ctx, table, pushc_prog, VX_MODE_GHOST, 1, settings.refinementRatio),
                   "limiter sync");
}

// ---------------------------------------------------------------- EMFieldSolver
EMFieldSolver::EMFieldSolver(Settings &s, const std::vector<std::shared_ptr<Mesh>> &m,
                             std::shared_ptr<Device> d)
    : settings(s), meshes(m), dev(std::move(d)) {
    // EMSolver.cpp:6-26, 88
    int N = (int)s.x_size_finest;
    fs.n_f = N;
    fs.n_prepad = std::max(std::floor(s.preLength / s.GetDx(0)), 2.0);
    fs.n_postpad = std::max(std::floor(s.postLength / s.GetDx(0)), 2.0);
    fs.n_e = N + fs.n_prepad + fs.n_postpad;
    fs.dx_f = s.GetDx(0);
    fs.dx2 = std::pow(s.GetDx(0), 2.0);
    fs.field_coef = (1.0 / (12 * s.GetDx(0)));
    fs.charge = alloc(N);
    fs.J = alloc(N);
    fs.neutral = alloc(N);
    fs.phi = alloc(N);
    fs.E = alloc(N);
    fs.Ex0 = alloc(1);
    fs.By = alloc((size_t)8 * fs.n_e);
    fs.Bz = alloc((size_t)8 * fs.n_e);
    fs.Ey = alloc((size_t)8 * fs.n_e);
    fs.Ez = alloc((size_t)8 * fs.n_e);
    fs.Ay = alloc((size_t)8 * fs.n_e);
    fs.Az = alloc((size_t)8 * fs.n_e);
    fs.a2 = alloc((size_t)N + 1);
    fs.work = alloc(vx_field_work_elems(N));
    charges_species = alloc((size_t)N * s.q.size());
    drive_d = alloc(12);
    double *E0 = alloc(N), *a20 = alloc((size_t)N + 1);
    fl.E0 = E0;
    fl.a20 = a20;
    fl.n_f = N;
    fl.n_prepad = fs.n_prepad;
    fl.n_e = fs.n_e;
    fl.E = fs.E;
    fl.a2 = fs.a2;
    fl.Ay1 = fs.Ay + fs.n_e;  // RK slot 1
    fl.Az1 = fs.Az + fs.n_e;
}

EMFieldSolver::~EMFieldSolver() {
    vx_moment_plan_destroy(plan);
    for (void *p : allocations) vx_free(dev->ctx, p);
}

double *EMFieldSolver::alloc(size_t n) {
    void *p = nullptr;
    dev->check(vx_malloc(dev->ctx, sizeof(double) * n, &p), "alloc field");
    allocations.push_back(p);
    return (double *)p;
}

// patch order of the reference's assembly: species, levels finest first,
// rectangles in level order (EMSolver.cpp:112-114, Mesh.cpp:52-56, Level.cpp:52-61)
void EMFieldSolver::RebuildMomentPlan() {
    vx_moment_plan_destroy(plan);
    plan = nullptr;
    std::vector<vx_patch> patches;
    std::vector<int32_t> starts;
    int nl = 0;
    for (auto &m : meshes) nl = std::max(nl, (int)m->levels.size());
    for (auto &m : meshes)
        for (int l = 0; l < nl; l++) {
            starts.push_back((int32_t)patches.size());
            if (l < (int)m->levels.size())
                for (auto &r : m->levels[l]->rectangles) patches.push_back(r->desc());
        }
    starts.push_back((int32_t)patches.size());
    vx_moment_layout lay{(int32_t)meshes.size(), nl, starts.data()};
    dev->check(vx_moment_plan_create(dev->ctx, patches.data(), &lay, fs.n_f, &plan), "moment plan");
}

// EMSolver.cpp:104-122
void EMFieldSolver::AssembleRhoAndJ() {
    for (auto &m : meshes)
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) r->CalculateRhoAndJ();
    dev->check(vx_moments_assemble(dev->ctx, plan, fs.charge, fs.J, charges_species), "assemble");
}

void EMFieldSolver::UpdatePotential() { dev->check(vx_poisson(dev->ctx, &fs), "poisson"); }

void EMFieldSolver::SnapshotStage0Fields() {
    dev->check(vx_memcpy_d2d(dev->ctx, (void *)fl.E0, fs.E, sizeof(double) * fs.n_f), "E0");
    dev->check(vx_memcpy_d2d(dev->ctx, (void *)fl.a20, fs.a2, sizeof(double) * (fs.n_f + 1)), "a2_0");
}

// EMSolver.cpp:194-202; the hooks are evaluated at x = 0 and the stage time
void EMFieldSolver::RGKStep(int step, double timestep) {
    if (device_drive) {
        dev->check(vx_maxwell_stage_dev(dev->ctx, &fs, step, timestep, drive_d + 2 * step), "maxwell");
        return;
    }
    double by0 = settings.GetBY(0, settings.time);
    double bz0 = settings.GetBZ(0, settings.time);
    dev->check(vx_maxwell_stage(dev->ctx, &fs, step, timestep, by0, bz0), "maxwell");
}

// EMSolver.cpp:621-629
void EMFieldSolver::EnforceChargeNeutralization() {
    AssembleRhoAndJ();
    dev->check(vx_neutralize(dev->ctx, &fs), "neutralize");
}

double EMFieldSolver::EstimateCFLBound() {
    std::vector<double> dpf;
    for (unsigned i = 0; i < settings.q.size(); i++) dpf.push_back(settings.GetDp(0, i));
    double out = 0;
    dev->check(vx_cfl_bound(dev->ctx, &fs, (int)settings.q.size(), settings.q.data(),
                            settings.m.data(), dpf.data(), &out),
               "cfl");
    return out;
}

// ---------------------------------------------------------------- SolverManager
// SolverManager.cpp:8-26
SolverManager::SolverManager(Settings &s, int device)
    : dev(std::make_shared<Device>(device)), settings(s) {
    {
        std::ostringstream o;  // SolverManager.cpp:9
        o << std::setfill('-') << std::setw(3) << '-' << " Time (T) " << std::setw(4) << '-' << '|'
          << std::setw(3) << '-' << " Particle " << std::setw(3) << '-' << '|' << std::setw(4) << '-'
          << " Levels " << std::setw(4) << '-' << '|' << std::setw(8) << '-' << " Rectangles "
          << std::setw(9) << '-' << "\n";
        ScreenWrite(o.str());
    }
    int numSpecies = (int)s.p_size.size();
    for (int i = 0; i < numSpecies; i++) meshes.push_back(std::make_shared<Mesh>(i, s, dev));
    EMSolver = std::make_shared<EMFieldSolver>(s, meshes, dev);
    {
        std::ostringstream o;  // SolverManager.cpp:17 (the time is written as 0)
        o << std::setfill(' ') << std::setw(4) << ' ' << std::fixed << 0.0 << std::setw(5) << ' ';
        ScreenWrite(o.str());
    }
    for (auto &m : meshes) {
        m->SetFieldSolver(EMSolver.get());
        m->PushData();
        ScreenRow(*m);
    }
    EMSolver->RebuildMomentPlan();
    EMSolver->EnforceChargeNeutralization();
}

void SolverManager::ScreenWrite(const std::string &s) {
    screen_ += s;
    if (echo_screen) {
        fputs(s.c_str(), stdout);
        fflush(stdout);
    }
}

// SolverManager.cpp:59-100: species, "used (allocated)" levels, patches per used level
void SolverManager::ScreenRow(const Mesh &mesh) {
    auto centred = [](const std::string &t, int target) {
        int pad = target - (int)t.size();
        return pad > 0 ? std::string(pad / 2, ' ') + t + std::string(pad / 2, ' ') : t;
    };
    std::ostringstream o, nums;
    if (mesh.particleType > 0) o << std::setw(17) << ' ';
    o << std::setw(16) << centred(std::to_string(mesh.particleType), 15) << std::setw(1) << ' ';
    int full = 0;
    for (auto &lvl : mesh.levels)
        if (!lvl->rectangles.empty()) {
            full++;
            nums << lvl->rectangles.size() << ", ";
        }
    std::string n = nums.str();
    if (n.size() >= 2) n.erase(n.size() - 2, 2);
    std::string lv = std::to_string(full) + " (" + std::to_string(mesh.levels.size()) + ")";
    o << std::setw(16) << centred(lv, 15) << std::setw(1) << ' ';
    o << std::setw(30) << centred(n, 29) << "\n";
    ScreenWrite(o.str());
}

// SolverManager.cpp:28-39 (SolverManager::Advance, host/graph.cpp, replays
// these launches as a graph)
void SolverManager::StepKernels(double timeStep) {
    for (int i = 0; i < 6; i++) {
        EMSolver->AssembleRhoAndJ();
        EMSolver->UpdatePotential();
        if (i == 0) EMSolver->SnapshotStage0Fields();
        for (auto &m : meshes) m->Advance(timeStep, i);
        settings.UpdateTime(i, timeStep);
        EMSolver->RGKStep(i, timeStep);
    }
}

// SolverManager.cpp:41-46
void SolverManager::AdvanceFields(double timeStep) {
    for (int i = 0; i < 6; i++) {
        settings.UpdateTime(i, timeStep);
        EMSolver->RGKStep(i, timeStep);
    }
}

void SolverManager::StageBegin(double dt, int step) {
    (void)dt;
    EMSolver->AssembleRhoAndJ();
    EMSolver->UpdatePotential();
    if (step == 0) EMSolver->SnapshotStage0Fields();
}

bool SolverManager::FusedItem(double dt, int step, vx_stage_item *item) {
    if (meshes.size() != 1) return false;
    Mesh &m = *meshes[0];
    if (!m.fused || !m.SinglePatch()) return false;
    Rectangle *r = m.levels.back()->rectangles[0].get();
    if (!r->AllWalls()) return false;
    r->FusedItem(dt, step, m.mixed, item);
    return true;
}

void SolverManager::StageEnd(double dt, int step) {
    settings.UpdateTime(step, dt);
    EMSolver->RGKStep(step, dt);
}

void SolverManager::CflAsync() {
    std::vector<double> dpf;
    for (unsigned i = 0; i < settings.q.size(); i++) dpf.push_back(settings.GetDp(0, i));
    dev->check(vx_cfl_bound_async(dev->ctx, &EMSolver->state(), (int)settings.q.size(), settings.q.data(),
                                  settings.m.data(), dpf.data()),
               "cfl async");
}

bool SolverManager::CflReady() {
    int r = 0;
    dev->check(vx_cfl_ready(dev->ctx, &r), "cfl ready");
    return r != 0;
}

double SolverManager::CflResult(double cfl) {
    double b = 0.0;
    dev->check(vx_cfl_result(dev->ctx, &b), "cfl result");
    return cfl * b;
}

void SolverManager::StageVlasov(double dt, int step) {
    if (step == 0) EMSolver->SnapshotStage0Fields();
    for (auto &m : meshes) m->Advance(dt, step);
}

void SolverManager::StageVlasovSub(double dt, int step, int sub) {
    for (auto &m : meshes) {
        if (sub == 0) m->Sub0(dt, step);
        else
            for (int i = (int)m->levels.size() - 1; i > -1; i--) m->levels[i]->FCTTimeStep(dt, step, sub);
        if (sub == 0) m->PushData(2);
        if (sub == 1) m->PushBoundaryC();
        if (sub == 2) m->PushData(1);
    }
}

void SolverManager::StageFields(double dt, int step) {
    settings.UpdateTime(step, dt);
    EMSolver->RGKStep(step, dt);
}

void SolverManager::SetFused(bool on) {
    DropGraphs();
    for (auto &m : meshes) m->fused = on;
}

void SolverManager::SetBatched(bool on) {
    DropGraphs();
    for (auto &m : meshes) m->batched = on;
}

void SolverManager::SetMixed(bool on) {
    DropGraphs();
    for (auto &m : meshes) m->mixed = on;
    EMSolver->compensated_moments = on;
}

}  // namespace vxh
