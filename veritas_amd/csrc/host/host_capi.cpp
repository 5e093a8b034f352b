// C entry points of the host library, used by the Python bindings
// (veritas_amd/__init__.py), the tests and bench.py.  A handle is one
// SolverManager for a scenario on one GPU.
#include <cstring>
#include <string>

#include "serve.hpp"
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

// f slot named "f0", "f1" or "f2", else -1
int slot_of(const std::string &w) { return w.size() == 2 && w[0] == 'f' && w[1] >= '0' && w[1] <= '2' ? w[1] - '0' : -1; }

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

// A run of the reference's inputs (Input, Particles as vx_settings_desc) and
// user hooks (vx_hooks); `hooks->user` must stay valid while the run lives.
void *vxh_create_desc(const vx_settings_desc *desc, const vx_hooks *hooks, int device) {
    try {
        Run *r = new Run;
        r->settings.reset(new vxh::Settings(*desc, *hooks));
        r->sm.reset(new vxh::SolverManager(*r->settings, device));
        return r;
    } catch (std::exception &e) {
        g_create_error = e.what();
        return nullptr;
    }
}

// Text of the run banner (Settings.cpp:60-125) or of the progress table
// (SolverManager.cpp:9-100): `which` 0 or 1.  Returns the length; copies when
// n is large enough (with a terminating 0).
int vxh_screen(void *h, int which, char *out, int n) {
    Run *r = (Run *)h;
    std::string t = which == 0 ? r->settings->Banner() : r->sm->Screen();
    if (out && n > (int)t.size()) std::memcpy(out, t.c_str(), t.size() + 1);
    return (int)t.size();
}

// Settings alone (no device): the banner of Settings(desc, hooks), or -1 with
// vxh_create_error() when the inputs are rejected (Settings.cpp:35-50).
int vxh_settings_banner(const vx_settings_desc *desc, const vx_hooks *hooks, char *out, int n) {
    try {
        vxh::Settings s(*desc, *hooks);
        std::string t = s.Banner();
        if (out && n > (int)t.size()) std::memcpy(out, t.c_str(), t.size() + 1);
        return (int)t.size();
    } catch (std::exception &e) {
        g_create_error = e.what();
        return -1;
    }
}

// The host's clustering (Mesh::splitRectangle, Mesh.cpp:644-792) on box
// (x0, p0)-(x1, p1) with n flagged cells (x, p pairs), no device: 4 ints per
// resulting box, returns 4 * count.
int vxh_split_box(int x0, int p0, int x1, int p1, const int *flags, int n, double minEfficiency, int *out,
                  int nout) {
    std::vector<vxh::coords> fl;
    for (int k = 0; k < n; k++) fl.push_back({flags[2 * k], flags[2 * k + 1]});
    vxh::level boxes = vxh::SplitBox({{x0, p0}, {x1, p1}}, fl, minEfficiency);
    if (out && nout >= 4 * (int)boxes.size())
        for (size_t b = 0; b < boxes.size(); b++) {
            out[4 * b] = boxes[b].first.first;
            out[4 * b + 1] = boxes[b].first.second;
            out[4 * b + 2] = boxes[b].second.first;
            out[4 * b + 3] = boxes[b].second.second;
        }
    return 4 * (int)boxes.size();
}

const char *vxh_create_error() { return g_create_error.c_str(); }
const char *vxh_last_error(void *h) { return ((Run *)h)->err.c_str(); }
void vxh_destroy(void *h) { delete (Run *)h; }

int vxh_calc_dt(void *h, double cfl, double *out) {
    return guarded(h, [&](Run *r) { *out = r->sm->CalculateDt(cfl); });
}

int vxh_set_batched(void *h, int on) {
    return guarded(h, [&](Run *r) { r->sm->SetBatched(on != 0); });
}

int vxh_set_fused(void *h, int on) {
    return guarded(h, [&](Run *r) { r->sm->SetFused(on != 0); });
}

int vxh_set_mixed(void *h, int on) {
    return guarded(h, [&](Run *r) { r->sm->SetMixed(on != 0); });
}

int vxh_set_graphs(void *h, int on) {
    return guarded(h, [&](Run *r) { r->sm->SetGraphs(on != 0); });
}

// out[0] = steps replayed from recorded graphs, out[1] = graphs recorded
int vxh_graph_stats(void *h, int *out) {
    return guarded(h, [&](Run *r) {
        out[0] = r->sm->graph_replays;
        out[1] = r->sm->graph_records;
    });
}

// strip height chosen for the fused stage of an n_x x n_p patch (0: not yet chosen)
int vxh_tile_rows(void *h, int n_x, int n_p, int step) {
    Run *r = (Run *)h;
    return vx_stage_tile_rows(r->sm->dev->ctx, n_x, n_p, step, r->sm->meshes[0]->mixed ? 1 : 0);
}

// resident waves per SIMD of the fused stage's kernel form chosen for that shape (0: not yet chosen)
int vxh_tile_waves(void *h, int n_x, int n_p, int step) {
    Run *r = (Run *)h;
    return vx_stage_tile_waves(r->sm->dev->ctx, n_x, n_p, step, r->sm->meshes[0]->mixed ? 1 : 0);
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

// Patch layout of species sp, same encoding as the oracle's vxo_layout:
// 9 ints per patch (level with 0 the finest, x_pos, p_pos, n_x, n_p, up,
// down, left, right), levels finest first.
int vxh_layout(void *h, int sp, int *out, int n) {
    Run *r = (Run *)h;
    try {
        std::vector<int> v;
        auto &levels = r->sm->meshes.at(sp)->levels;
        for (unsigned l = 0; l < levels.size(); l++)
            for (auto &p : levels[l]->rectangles) {
                int e[9] = {(int)l, p->x_pos, p->p_pos, p->n_x, p->n_p, p->up, p->down, p->left, p->right};
                v.insert(v.end(), e, e + 9);
            }
        if (out && n >= (int)v.size()) std::memcpy(out, v.data(), sizeof(int) * v.size());
        return (int)v.size();
    } catch (std::exception &e) {
        r->err = e.what();
        return -1;
    }
}

// Plane of patch idx of level lvl (0 = finest) of species sp; -2 unknown name or patch
int vxh_patch_get(void *h, int sp, int lvl, int idx, const char *what, int ghosts, double *out, int n) {
    Run *r = (Run *)h;
    try {
        auto &levels = r->sm->meshes.at(sp)->levels;
        if (lvl < 0 || lvl >= (int)levels.size() || idx < 0 || idx >= (int)levels[lvl]->rectangles.size())
            return -2;
        vxh::Rectangle &p = *levels[lvl]->rectangles[idx];
        if (slot_of(what) >= 0) return copy_out(p.DownloadSlot(slot_of(what), ghosts), out, n);
        const double *pl = p.Plane(what);
        if (!pl) return -2;
        return copy_out(p.Download(pl, ghosts), out, n);
    } catch (std::exception &e) {
        r->err = e.what();
        return -1;
    }
}

// Refinement flags of hierarchy level lvl (0 = coarsest) from the current
// data, as (i, j) pairs; returns 2 * count
int vxh_flags(void *h, int sp, int lvl, int *out, int n) {
    Run *r = (Run *)h;
    try {
        std::vector<vxh::coords> fl;
        r->sm->meshes.at(sp)->getError(lvl, false, fl);
        if (out && n >= 2 * (int)fl.size())
            for (size_t k = 0; k < fl.size(); k++) {
                out[2 * k] = fl[k].first;
                out[2 * k + 1] = fl[k].second;
            }
        return 2 * (int)fl.size();
    } catch (std::exception &e) {
        r->err = e.what();
        return -1;
    }
}

// signatures kept by the last getError of patch idx of level lvl (0 = the
// coarsest, as vxh_flags): out = lo_x, lo_p, n_rows, n_cols, rows..., cols...;
// returns the length
int vxh_signatures(void *h, int sp, int lvl, int idx, int *out, int n) {
    Run *r = (Run *)h;
    try {
        auto &levels = r->sm->meshes.at(sp)->levels;
        const vxh::Rectangle &R = *levels.at(r->settings->maxDepth - lvl)->rectangles.at(idx);
        std::vector<int> v = {R.sig_lo_x, R.sig_lo_p, (int)R.sig_x.size(), (int)R.sig_p.size()};
        v.insert(v.end(), R.sig_x.begin(), R.sig_x.end());
        v.insert(v.end(), R.sig_p.begin(), R.sig_p.end());
        if (out && n >= (int)v.size()) std::memcpy(out, v.data(), sizeof(int) * v.size());
        return (int)v.size();
    } catch (std::exception &e) {
        r->err = e.what();
        return -1;
    }
}

// Task batches of species sp's programs: for each program (PushData's,
// PushBoundaryC's, then each level's coarse-fine fluxes) its batch count
// followed by the tasks in each batch.  Returns the number of ints.
int vxh_program_stats(void *h, int sp, int *out, int n) {
    Run *r = (Run *)h;
    std::vector<int> v;
    try {
        for (auto &b : r->sm->meshes.at(sp)->ProgramBatches()) {
            v.push_back((int)b.size());
            v.insert(v.end(), b.begin(), b.end());
        }
    } catch (std::exception &e) {
        r->err = e.what();
        return -1;
    }
    if (out && n >= (int)v.size()) std::memcpy(out, v.data(), sizeof(int) * v.size());
    return (int)v.size();
}

// SolverManager::reGrid(t) (SolverManager.cpp:48-58)
int vxh_regrid(void *h, double t) { return guarded(h, [&](Run *r) { r->sm->reGrid(t); }); }

// The nesting list of Mesh::updateHierarchy (host logic, for the CPU tests):
// level cells under the finer-level box [i0, i1] x [j0, j1]; writes up to n
// (ic, jc) pairs, returns the count
int vxh_cells_under_box(int nmax, int pmax, double fx, double fp, int i0, int i1, int j0, int j1, int *out, int n) {
    std::vector<vxh::coords> v;
    vxh::cells_under_box(nmax, pmax, fx, fp, i0, i1, j0, j1, v);
    const int m = (int)v.size();
    if (out && n >= m)
        for (int k = 0; k < m; k++) {
            out[2 * k] = v[k].first;
            out[2 * k + 1] = v[k].second;
        }
    return m;
}

// regrids of species sp's mesh since set-up, and how many of them rebuilt the patches
int vxh_regrid_stats(void *h, int sp, int *out) {
    return guarded(h, [&](Run *r) {
        const vxh::Mesh &m = *r->sm->meshes.at(sp);
        out[0] = m.regrids;
        out[1] = m.rebuilds;
    });
}

// SolverManager::fileOutput / OutputRectangles: text files under output/ of
// the working directory, written by a background thread; vxh_output_flush waits.
int vxh_file_output(void *h, double t) { return guarded(h, [&](Run *r) { r->sm->fileOutput(t); }); }
int vxh_output_rectangles(void *h, double t) { return guarded(h, [&](Run *r) { r->sm->OutputRectangles(t); }); }
int vxh_output_flush(void *h) { return guarded(h, [&](Run *r) { r->sm->FlushOutput(); }); }

int vxh_save(void *h, const char *path) { return guarded(h, [&](Run *r) { r->sm->SaveCheckpoint(path); }); }
int vxh_load(void *h, const char *path) { return guarded(h, [&](Run *r) { r->sm->LoadCheckpoint(path); }); }

// f slot "f0"/"f1"/"f2" of the base patch from n = (n_x+2g)(n_p+2g) values
// (ghost layers beyond g keep their contents); -2 unknown name or size
int vxh_put(void *h, const char *what, int ghosts, const double *in, int n) {
    Run *r = (Run *)h;
    int s = slot_of(what);
    vxh::Rectangle &p = base(r);
    if (s < 0 || ghosts < 0 || ghosts > 2 || (size_t)n != (size_t)(p.n_x + 2 * ghosts) * (p.n_p + 2 * ghosts))
        return -2;
    return guarded(h, [&](Run *) { p.UploadSlot(s, ghosts, in); });
}

// Named arrays, same names and layouts as the oracle's C API.
int vxh_get(void *h, const char *what, int ghosts, double *out, int n) {
    Run *r = (Run *)h;
    try {
        std::string w(what);
        vxh::Rectangle &p = base(r);
        if (slot_of(w) >= 0) return copy_out(p.DownloadSlot(slot_of(w), ghosts), out, n);
        if (const double *pl = p.Plane(w)) return copy_out(p.Download(pl, ghosts), out, n);
        const vx_field_state &fs = r->sm->EMSolver->state();
        int N = fs.n_f;
        size_t ne8 = (size_t)8 * fs.n_e;
        if (w == "phi") return copy_out(dl(r, fs.phi, N), out, n);
        if (w == "charge") return copy_out(dl(r, fs.charge, N), out, n);
        if (w.rfind("charge_sp", 0) == 0) {  // species i of the last assembly
            const int sp = std::stoi(w.substr(9));
            const double *cs = r->sm->EMSolver->charges_species;
            if (!cs || sp < 0 || sp >= (int)r->sm->meshes.size()) return -2;
            return copy_out(dl(r, cs + (size_t)sp * N, N), out, n);
        }
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
        if (w == "dNdP") {
            r->sm->EMSolver->AssembleEnergy();
            return copy_out(r->sm->EMSolver->energies.at(0), out, n);
        }
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


// ---- serving loop (host/serve.cpp) ----
struct ServeHandle {
    std::unique_ptr<vxh::Server> srv;
    std::string err;
};

void *vxh_server_create(int device, double budget_bytes, int max_active) {
    try {
        ServeHandle *h = new ServeHandle;
        h->srv.reset(new vxh::Server(device, (size_t)budget_bytes, max_active));
        return h;
    } catch (std::exception &e) {
        g_create_error = e.what();
        return nullptr;
    }
}

void vxh_server_destroy(void *h) { delete (ServeHandle *)h; }
const char *vxh_server_error(void *h) { return ((ServeHandle *)h)->err.c_str(); }

int vxh_server_submit(void *h, int id, const vx_scenario_params *prm, int steps, double cfl, double arrival,
                      int mixed) {
    vxh::ServeRequest r;
    r.id = id;
    r.params = *prm;
    r.steps = steps;
    r.cfl = cfl;
    r.arrival = arrival;
    r.mixed = mixed != 0;
    ((ServeHandle *)h)->srv->Submit(r);
    return 0;
}

int vxh_server_keep_state(void *h, int on) {
    ((ServeHandle *)h)->srv->keep_state = on != 0;
    return 0;
}

int vxh_server_run(void *h) {
    ServeHandle *s = (ServeHandle *)h;
    try {
        s->srv->Run();
        return 0;
    } catch (std::exception &e) {
        s->err = e.what();
        return -1;
    }
}

// 9 doubles per finished request: id, arrival, admitted, started, finished,
// footprint bytes, cell-updates, batched, simulated end time
int vxh_server_records(void *h, double *out, int n) {
    const auto &rec = ((ServeHandle *)h)->srv->records;
    int m = (int)rec.size() * 9;
    if (out && n >= m)
        for (size_t k = 0; k < rec.size(); k++) {
            const auto &r = rec[k];
            double v[9] = {(double)r.id, r.arrival, r.admitted, r.started, r.finished, (double)r.footprint,
                           r.cell_updates, r.batched ? 1.0 : 0.0, r.time};
            std::memcpy(out + 9 * k, v, sizeof v);
        }
    return m;
}

// peak footprint bytes, peak active runs, ticks, batched launches
int vxh_server_stats(void *h, double *out) {
    const auto &s = *((ServeHandle *)h)->srv;
    out[0] = (double)s.peak_bytes;
    out[1] = s.peak_active;
    out[2] = (double)s.ticks;
    out[3] = (double)s.batched_launches;
    return 0;
}

int vxh_server_final_f(void *h, int id, double *out, int n) {
    const auto &m = ((ServeHandle *)h)->srv->final_f;
    auto it = m.find(id);
    if (it == m.end()) return -1;
    return copy_out(it->second, out, n);
}

}  // extern "C"
