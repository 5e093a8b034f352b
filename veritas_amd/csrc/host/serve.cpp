// Serving loop (continuous batching of independent runs on one GPU).
//
// Requests arrive on the wall clock.  A request is admitted when it has
// arrived, a slot is free (at most max_active runs at once) and its estimated
// device footprint fits the budget beside the runs already held; a worker
// thread then builds it (settings, initial distribution, uploads, the first
// dt) while the main thread keeps stepping the others.  Runs progress
// independently: each pass of the main loop takes the runs whose previous step
// has finished (their CFL bound has arrived; polled, not waited for) and
// advances each of them by one step, stage by stage in lock-step:
//   * each run's moments, Poisson solve and stage-0 snapshot on its own stream;
//   * the Vlasov stage of every single walled patch in one shared launch on
//     the server's stream (vx_fct_stage_fused_batch), ordered after the runs'
//     streams and before them again with stream waits; adaptive-mesh runs
//     take their own sub-step kernels on their own streams;
//   * each run's stage clock and Maxwell stage on its own stream.
// After its step every run queues its CFL bound on the device
// (vx_cfl_bound_async); a small run thus takes many steps while a large one
// takes one.  A finished run is destroyed at once, which frees its memory for
// the next admission.
//
// Each run computes exactly what it computes alone: the same kernels with the
// same arguments, and the shared launch gives each patch the bits of its own
// launch; dt is cfl x CalculateDt before every step, as the reference's driver
// takes it (veritas.cpp:131-137).
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace vxh {

namespace {
double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

size_t cells_of(const SolverManager &sm) {
    size_t c = 0;
    for (auto &m : sm.meshes)
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) c += (size_t)r->n_x * r->n_p;
    return c;
}
}  // namespace

// Device bytes a run will hold: 23 planes per patch cell (either Vlasov path
// can run) and the field arrays; a run with refinement levels is charged as
// if each finer level held twice the base level's cells (ratio-2 levels
// covering half of the coarser one), then charged its measured bytes once
// built.
size_t ServeEstimate(const vx_scenario_params &p) {
    const size_t plane = vx_plane_elems(p.nx, p.np) * sizeof(double);
    const int levels = std::max(1, (int)p.levels);
    size_t patches = 23 * plane * (size_t)(1 + 2 * (levels - 1));
    const size_t nf = (size_t)p.nx * (size_t)std::pow((double)std::max(1, (int)p.r), levels - 1);
    const size_t fields = sizeof(double) * (24 * nf + 48 * (nf + 8));
    return patches + fields;
}

Server::Server(int dev, size_t budget_, int max_active_)
    : device(dev), budget(budget_), max_active(std::max(1, std::min(max_active_, 64))) {
    batch.reset(new Device(device));
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto &w : workers)
        if (w.joinable()) w.join();
}

void Server::Submit(const ServeRequest &r) {
    auto it = std::upper_bound(pending.begin(), pending.end(), r,
                               [](const ServeRequest &a, const ServeRequest &b) { return a.arrival < b.arrival; });
    pending.insert(it, r);
}

void Server::Build() {
    for (;;) {
        std::unique_ptr<Live> L;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [this] { return stop || !to_build.empty(); });
            if (stop) return;
            L = std::move(to_build.front());
            to_build.pop_front();
        }
        try {
            vx::Scenario scn = vx::Scenario::from_params(L->req.params);
            L->settings.reset(new Settings(scn));
            L->sm.reset(new SolverManager(*L->settings, device));
            L->sm->SetGraphs(false);  // the loop drives the stages itself
            if (L->req.mixed) L->sm->SetMixed(true);
            L->dt = L->sm->CalculateDt(L->req.cfl);
            L->bytes = vx_ctx_live_bytes(L->sm->dev->ctx);
            L->rec.cell_updates = (double)cells_of(*L->sm) * L->req.steps;
        } catch (std::exception &e) {
            L->error = e.what();
        }
        {
            std::lock_guard<std::mutex> g(mu);
            built.push_back(std::move(L));
        }
        cv.notify_all();
    }
}

void Server::Tick(const std::vector<Live *> &ready) {
    std::vector<vx_stage_item> items;
    std::vector<Live *> shared;
    for (int s = 0; s < 6; s++) {
        items.clear();
        shared.clear();
        for (Live *L : ready) L->sm->StageBegin(L->dt, s);
        for (Live *L : ready) {
            vx_stage_item it;
            if (L->sm->FusedItem(L->dt, s, &it)) {
                items.push_back(it);
                shared.push_back(L);
            } else {
                for (auto &m : L->sm->meshes) m->Advance(L->dt, s);
            }
        }
        if (!items.empty()) {
            for (Live *L : shared) batch->check(vx_ctx_wait(batch->ctx, L->sm->dev->ctx), "batch waits for run");
            batch->check(vx_fct_stage_fused_batch(batch->ctx, items.data(), (int)items.size(), s), "stage batch");
            batched_launches++;
            for (Live *L : shared) batch->check(vx_ctx_wait(L->sm->dev->ctx, batch->ctx), "run waits for batch");
        }
        for (Live *L : ready) L->sm->StageEnd(L->dt, s);
    }
    for (Live *L : ready) {
        L->done++;
        L->rec.batched = L->rec.batched || std::find(shared.begin(), shared.end(), L) != shared.end();
        L->sm->CflAsync();  // next dt, or the end-of-run marker
        L->waiting = true;
    }
    ticks++;
}

void Server::Run() {
    const auto t0 = std::chrono::steady_clock::now();
    while ((int)workers.size() < std::max(1, builders)) workers.emplace_back([this] { Build(); });
    std::vector<std::unique_ptr<Live>> active;
    size_t held = 0;   // footprints of admitted runs (estimates until built)
    int building = 0;
    std::string failure;
    while (!pending.empty() || building > 0 || !active.empty()) {
        // admissions, in arrival order
        while (!pending.empty() && pending.front().arrival <= now_s(t0) &&
               (int)active.size() + building < max_active) {
            const size_t est = ServeEstimate(pending.front().params);
            if (held + est > budget && (held > 0 || building > 0)) break;  // wait for memory
            auto L = std::make_unique<Live>();
            L->req = pending.front();
            pending.pop_front();
            L->rec.id = L->req.id;
            L->rec.arrival = L->req.arrival;
            L->rec.admitted = now_s(t0);
            L->bytes = est;
            held += est;
            peak_bytes = std::max(peak_bytes, held);
            building++;
            {
                std::lock_guard<std::mutex> g(mu);
                to_build.push_back(std::move(L));
            }
            cv.notify_all();
        }
        // runs the worker has built
        {
            std::lock_guard<std::mutex> g(mu);
            while (!built.empty()) {
                auto L = std::move(built.front());
                built.pop_front();
                building--;
                held -= ServeEstimate(L->req.params);
                if (!L->error.empty()) {
                    if (failure.empty()) failure = "request " + std::to_string(L->req.id) + ": " + L->error;
                    continue;
                }
                held += L->bytes;
                peak_bytes = std::max(peak_bytes, held);
                L->rec.started = now_s(t0);
                L->rec.footprint = L->bytes;
                active.push_back(std::move(L));
            }
        }
        if (!failure.empty()) break;
        if (active.empty()) {
            std::unique_lock<std::mutex> lk(mu);
            double wait = pending.empty() ? 0.05 : std::max(0.0, pending.front().arrival - now_s(t0));
            cv.wait_for(lk, std::chrono::duration<double>(std::min(wait, 0.05)));
            continue;
        }
        peak_active = std::max(peak_active, (int)active.size());
        // runs whose last step has finished: they leave, or take their next dt
        std::vector<Live *> ready;
        for (size_t k = 0; k < active.size();) {
            Live &L = *active[k];
            if (L.waiting && !L.sm->CflReady()) {
                k++;
                continue;
            }
            if (L.waiting) {
                L.dt = L.sm->CflResult(L.req.cfl);
                L.waiting = false;
            }
            if (L.done >= L.req.steps) {
                L.sm->dev->check(vx_sync(L.sm->dev->ctx), "finish");
                L.rec.finished = now_s(t0);
                L.rec.time = L.settings->time;
                if (keep_state) {
                    const Rectangle &r = *L.sm->meshes[0]->levels.back()->rectangles[0];
                    final_f[L.req.id] = r.DownloadSlot(1, 0);
                }
                records.push_back(L.rec);
                held -= L.bytes;
                active.erase(active.begin() + k);
                continue;
            }
            ready.push_back(&L);
            k++;
        }
        if (ready.empty()) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
        }
        Tick(ready);
    }
    if (!failure.empty()) {
        // a run could not be built: admit nothing more, take back the builds
        // not yet started, wait for those under way, and drop the active runs,
        // so that the server is left empty (and usable) before reporting
        pending.clear();
        std::unique_lock<std::mutex> lk(mu);
        building -= (int)to_build.size();
        to_build.clear();
        while (building > 0) {
            cv.wait(lk, [this] { return !built.empty(); });
            while (!built.empty()) {
                built.pop_front();
                building--;
            }
        }
        lk.unlock();
        for (auto &L : active)
            if (L->sm) (void)vx_sync(L->sm->dev->ctx);
        active.clear();
        throw Error("serve: " + failure);
    }
}

}  // namespace vxh
