// SolverManager::Advance (SolverManager.cpp:28-39) as a graph launch.
//
// A step's launches depend on the host state only through the plane pointers
// of every patch, the device tables, dt and the hook values.  So the launches
// of one step from a given state are recorded once (vx_graph_begin/end) and
// replayed for every later step from the same state; the hook values of the
// six stages go to device memory before each replay.  The first step from a
// state runs eagerly (it does the one-off work: table uploads, program checks,
// buffer growth, none of which can be recorded); the second records.  A
// recording that fails (some entry point needed to synchronise) leaves that
// state to eager steps.  The fused single-patch path cycles its three f planes
// with period two after the first step, so it settles on two graphs.
#include <cstring>

#include "veritas.hpp"

namespace vxh {

void SolverManager::SetGraphs(bool on) {
    graphs_on = on;
    if (!on) DropGraphs();
}

void SolverManager::DropGraphs() {
    step_graphs.clear();
    graph_seen.clear();
    graph_refused.clear();
}

std::vector<uintptr_t> SolverManager::GraphKey(double dt) const {
    std::vector<uintptr_t> k;
    uint64_t bits;
    std::memcpy(&bits, &dt, sizeof bits);
    k.push_back((uintptr_t)bits);
    for (auto &m : meshes) {
        k.push_back((uintptr_t)m->fused);
        k.push_back((uintptr_t)m->table_version);
        for (auto &l : m->levels) {
            k.push_back(l->rectangles.size());
            for (auto &r : l->rectangles) {
                const vx_patch &P = r->desc();
                for (int s = 0; s < 3; s++) k.push_back((uintptr_t)P.f[s]);
            }
        }
    }
    return k;
}

SolverManager::HostState SolverManager::SaveHostState() const {
    HostState h;
    for (auto &m : meshes) {
        h.table_versions.push_back(m->table_version);
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) h.slots.push_back(r->GetSlots());
    }
    h.time = settings.time;
    return h;
}

void SolverManager::RestoreHostState(const HostState &h) {
    size_t q = 0, t = 0;
    for (auto &m : meshes) {
        m->table_version = h.table_versions.at(t++);
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) r->SetSlots(h.slots.at(q++));
    }
}

SolverManager::StepGraph *SolverManager::RecordStep(double dt, const std::vector<uintptr_t> &key) {
    const HostState pre = SaveHostState();
    EMSolver->device_drive = true;
    dev->check(vx_graph_begin(dev->ctx), "graph begin");
    vx_graph *g = nullptr;
    bool ok = true;
    try {
        StepKernels(dt);
    } catch (const Error &) {
        ok = false;
    }
    const HostState post = SaveHostState();
    int rc = vx_graph_end(dev->ctx, ok ? &g : nullptr);
    EMSolver->device_drive = false;
    RestoreHostState(pre);
    settings.time = pre.time;
    if (!ok || rc != VX_OK || !g) {
        vx_graph_destroy(g);
        if (graph_refused.size() >= 16) graph_refused.erase(graph_refused.begin());
        graph_refused.push_back(key);
        return nullptr;
    }
    if (step_graphs.size() >= 8) step_graphs.erase(step_graphs.begin());
    auto sg = std::make_unique<StepGraph>();
    sg->key = key;
    sg->post = post;
    sg->g = g;
    step_graphs.push_back(std::move(sg));
    graph_records++;
    return step_graphs.back().get();
}

void SolverManager::Advance(double timeStep) {
    if (!graphs_on) {
        StepKernels(timeStep);
        return;
    }
    const std::vector<uintptr_t> key = GraphKey(timeStep);
    StepGraph *sg = nullptr;
    for (auto &g : step_graphs)
        if (g->key == key) sg = g.get();
    if (!sg) {
        bool seen = false, refused = false;
        for (auto &k : graph_seen) seen = seen || k == key;
        for (auto &k : graph_refused) refused = refused || k == key;
        if (!seen || refused) {
            if (!seen) {
                // a caller that changes dt every step (the reference's driver
                // takes min(CalculateDt, dt_max), veritas.cpp:131-137) never
                // returns to a state; keep only the latest keys
                if (graph_seen.size() >= 16) graph_seen.erase(graph_seen.begin());
                graph_seen.push_back(key);
            }
            StepKernels(timeStep);
            return;
        }
        sg = RecordStep(timeStep, key);
        if (!sg) {
            StepKernels(timeStep);
            return;
        }
    }
    // the stage clock and the hooks at x = 0, as the eager step evaluates them
    // (Settings.cpp:166-179 after each stage's Vlasov update, EMSolver.cpp:492-499)
    double drive[12];
    for (int i = 0; i < 6; i++) {
        settings.UpdateTime(i, timeStep);
        drive[2 * i] = settings.GetBY(0, settings.time);
        drive[2 * i + 1] = settings.GetBZ(0, settings.time);
    }
    dev->check(vx_upload_small(dev->ctx, EMSolver->drive_d, drive, sizeof drive), "hook values");
    dev->check(vx_graph_launch(dev->ctx, sg->g), "graph launch");
    RestoreHostState(sg->post);
    graph_replays++;
}

}  // namespace vxh
