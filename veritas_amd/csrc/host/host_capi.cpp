// C entry points of the host library, used by the Python bindings
// (veritas_amd/__init__.py), the tests and bench.py.  A handle is one
// SolverManager for a scenario on one GPU.
#include <cstring>
#include <string>

#include "veritas.hpp"

namespace {
struct Run {
    std::unique_ptr<vxh::Settings> settings;
    std::unique_ptr<vxh::SolverManager> sm;
    std::string err;
};
thread_local std::string g_create_error;

int copy_out(const std::vector<double> &v, double *out, int n) {
    int m = (int)v.size();
    if (out && n >= m) std::memcpy(out, v.data(), sizeof(double) * m);
    return m;
}

vxh::Rectangle &base(Run *r) { return *r->sm->meshes[0]->levels.back()->rectangles[0]; }

std::vector<double> dl(Run *r, const double *d, size_t n) {
    std::vector<double> v(n);
    r->sm->dev->check(vx_memcpy_d2h(r->sm->dev->ctx, v.data(), d, sizeof(double) * n), "download");
    return v;
}

template <class F>
int guarded(void *h, F f) {
    Run *r = (Run *)h;
    try {
        f(r);
        return 0;
    } catch (std::exception &e) {
        r->err = e.what();
        return -1;
    }
}
}  // namespace

extern "C" {

void *vxh_create(const vx_scenario_params *prm, int device) {
    try {
        Run *r = new Run;
        vx::Scenario scn = vx::Scenario::from_params(*prm);
        r->settings.reset(new vxh::Settings(scn));
        r->sm.reset(new vxh::SolverManager(*r->settings, device));
        return r;
    } catch (std::exception &e) {
        g_create_error = e.what();
        return nullptr;
    }
}

const char *vxh_create_error() { return g_create_error.c_str(); }
const char *vxh_last_error(void *h) { return ((Run *)h)->err.c_str(); }
void vxh_destroy(void *h) { delete (Run *)h; }

int vxh_calc_dt(void *h, double cfl, double *out) {
    return guarded(h, [&](Run *r) { *out = r->sm->CalculateDt(cfl); });
}

int vxh_set_fused(void *h, int on) {
    return guarded(h, [&](Run *r) { r->sm->SetFused(on != 0); });
}

int vxh_advance(void *h, double dt, int nsteps) {
    return guarded(h, [&](Run *r) {
        for (int i = 0; i < nsteps; i++) r->sm->Advance(dt);
    });
}

int vxh_advance_fields(void *h, double dt, int nsteps) {
    return guarded(h, [&](Run *r) {
        for (int i = 0; i < nsteps; i++) r->sm->AdvanceFields(dt);
    });
}

int vxh_sync(void *h) {
    return guarded(h, [&](Run *r) { r->sm->dev->check(vx_sync(r->sm->dev->ctx), "sync"); });
}

int vxh_stage_moments(void *h) { return guarded(h, [&](Run *r) { r->sm->StageMoments(); }); }
int vxh_stage_poisson(void *h) { return guarded(h, [&](Run *r) { r->sm->StagePoisson(); }); }
int vxh_stage_vlasov(void *h, double dt, int step) {
    return guarded(h, [&](Run *r) { r->sm->StageVlasov(dt, step); });
}
int vxh_stage_vlasov_sub(void *h, double dt, int step, int sub) {
    return guarded(h, [&](Run *r) { r->sm->StageVlasovSub(dt, step, sub); });
}
int vxh_stage_fields(void *h, double dt, int step) {
    return guarded(h, [&](Run *r) { r->sm->StageFields(dt, step); });
}

// Run `warmup` untimed steps, then `steps` steps bracketed by HIP events on
// the library stream; *ms receives the elapsed device time.
int vxh_time_advance(void *h, double dt, int warmup, int steps, float *ms) {
    return guarded(h, [&](Run *r) {
        vx_ctx *c = r->sm->dev->ctx;
        for (int i = 0; i < warmup; i++) r->sm->Advance(dt);
        r->sm->dev->check(vx_sync(c), "sync");
        r->sm->dev->check(vx_event_record(c, 0), "event");
        for (int i = 0; i < steps; i++) r->sm->Advance(dt);
        r->sm->dev->check(vx_event_record(c, 1), "event");
        r->sm->dev->check(vx_event_elapsed_ms(c, 0, 1, ms), "elapsed");
    });
}

int vxh_timing_begin(void *h, const char *kernel) {
    return guarded(h, [&](Run *r) { r->sm->dev->check(vx_timing_begin(r->sm->dev->ctx, kernel), "timing"); });
}

int vxh_timing_end(void *h, double *total_ms, int *launches) {
    return guarded(h, [&](Run *r) {
        r->sm->dev->check(vx_timing_end(r->sm->dev->ctx, total_ms, launches), "timing");
    });
}

// Named arrays, same names and layouts as the oracle's C API.
int vxh_get(void *h, const char *what, int ghosts, double *out, int n) {
    Run *r = (Run *)h;
    try {
        std::string w(what);
        vxh::Rectangle &p = base(r);
        if (const double *pl = p.Plane(w)) return copy_out(p.Download(pl, ghosts), out, n);
        const vx_field_state &fs = r->sm->EMSolver->state();
        int N = fs.n_f;
        size_t ne8 = (size_t)8 * fs.n_e;
        if (w == "phi") return copy_out(dl(r, fs.phi, N), out, n);
        if (w == "charge") return copy_out(dl(r, fs.charge, N), out, n);
        if (w == "J") return copy_out(dl(r, fs.J, N), out, n);
        if (w == "neutral") return copy_out(dl(r, fs.neutral, N), out, n);
        if (w == "a2") return copy_out(dl(r, fs.a2, N + 1), out, n);
        if (w == "E") return copy_out(dl(r, fs.E, N), out, n);
        if (w == "By") return copy_out(dl(r, fs.By, ne8), out, n);
        if (w == "Bz") return copy_out(dl(r, fs.Bz, ne8), out, n);
        if (w == "Ey") return copy_out(dl(r, fs.Ey, ne8), out, n);
        if (w == "Ez") return copy_out(dl(r, fs.Ez, ne8), out, n);
        if (w == "Ay") return copy_out(dl(r, fs.Ay, ne8), out, n);
        if (w == "Az") return copy_out(dl(r, fs.Az, ne8), out, n);
        if (w == "scalars") {
            const vx_patch &d = p.desc();
            std::vector<double> ex0 = dl(r, fs.Ex0, 1);
            std::vector<double> s = {ex0[0], r->settings->time, d.dx, d.dp, d.pmin, d.m, d.q,
                                     (double)fs.n_prepad};
            return copy_out(s, out, n);
        }
        return -2;
    } catch (std::exception &e) {
        r->err = e.what();
        return -1;
    }
}

}  // extern "C"
