// Host-side classes of the MI355X framework.  They keep the reference's class
// interface (SolverManager.hpp, Mesh.hpp, Level.hpp, Rectangle.hpp,
// EMSolver.hpp, BoundaryCondition.hpp, Settings.hpp) so a driver written
// against the reference reads the same, while all numerics run in the HIP
// kernels behind the C ABI of include/veritas_amd.h.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/veritas_amd.h"
#include "../scenario.h"

namespace vxh {

class Rectangle;
class Level;
class Mesh;
class EMFieldSolver;
class SolverManager;

// thrown when a C-ABI call fails (the reference exits; a library reports)
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Settings.hpp:4-31
struct Input {
    double minEfficiency = 0.6, dx = 0.5, k = 0.01;
    double refinementCriteria = 1e-8, cfl = 0.9, sizeWeight = 0.0;
    double preLength = 2, postLength = 2;
    unsigned int nx = 76 * 4, r = 2, Lfinest = 5;
    std::vector<double> tempEM = {0.0};
    // Settings::quadratureDepth before settingsOverride runs (the reference
    // leaves it to the override, veritas.cpp:51)
    int quadratureDepth = 1;
};
struct Particles {
    std::vector<std::vector<double>> misc = {{0.0, 0.01}};  // Settings::temp, for the user's hooks
    std::vector<double> mass = {9.10938291e-31};
    std::vector<double> charge = {-1.60217657e-19};
    std::vector<unsigned int> np = {50};
    std::vector<double> dp = {0.1};
    std::vector<double> pmin = {0.1};
};
struct Output {
    bool time = true, rectangleData = true, charge = true, energy = true, potential = true;
    bool EFieldLongitudinal = true, EFieldTransverse = true, BFieldTransverse = true,
         AFieldSquared = true;
    int precision = 15;
};

class Settings;

// The user hooks of Settings.hpp:45-64 (written by the user in
// veritas.cpp:36-115) as callables.  GetBY/GetBZ left empty drive nothing,
// an empty RefinementOverride never flags, an empty settingsOverride changes
// nothing; InitialDistribution is required.
struct Hooks {
    std::function<double(double x, double p, int particleType)> InitialDistribution;
    std::function<double(double x, double t)> GetBY, GetBZ;
    std::function<bool(double x, double p, int depth, int particleType)> RefinementOverride;
    std::function<void(Settings &)> settingsOverride;
    // the same hooks from C function pointers (vx_hooks)
    static Hooks FromC(const vx_hooks &h);
};

// Settings.hpp:33-65
class Settings {
public:
    Output output;
    double dx, time, minEfficiency, preLength, postLength, refinementCriteria, cfl, sizeWeight;
    std::vector<double> m, q, dp, fMax, pmin;
    unsigned int x_size_finest, x_size, refinementRatio;
    int maxDepth, quadratureDepth = 1;
    std::vector<unsigned int> p_size, p_size_finest;
    std::vector<double> tempEM;               // Input::tempEM, for the user's hooks
    std::vector<std::vector<double>> temp;    // Particles::misc, for the user's hooks
    Hooks hooks;
    // the physics of a built-in scenario behind the hooks (empty for user
    // hooks), which checkpoints compare
    std::vector<double> identity;

    // Settings.cpp:5-58: copy the inputs, run settingsOverride, check, derive
    Settings(const Input &grid, const Particles &particles, const Output &out, Hooks hooks);
    // the same from the C description (vx_settings_desc) and C hooks
    Settings(const vx_settings_desc &desc, const vx_hooks &hooks);
    // a built-in scenario (scenario.h) as the provider of inputs and hooks
    explicit Settings(const vx::Scenario &scenario);
    double GetDp(int level, int particleType) const;
    double GetDx(int level) const;
    int GetXSize(int level) const;
    int GetPSize(int level, int particleType) const;
    double GetMass(int i) const { return m.at(i); }
    double GetCharge(int i) const { return q.at(i); }
    double GetfMax(int i) const { return fMax.at(i); }
    void UpdateTime(int step, double dt);
    void DetermineMaximum();
    double InitialDistribution(double x, double p, int particleType) const {
        return hooks.InitialDistribution(x, p, particleType);
    }
    double GetBY(double x, double t) const { return hooks.GetBY ? hooks.GetBY(x, t) : 0.0; }
    double GetBZ(double x, double t) const { return hooks.GetBZ ? hooks.GetBZ(x, t) : 0.0; }
    bool HasRefinementOverride() const { return (bool)hooks.RefinementOverride; }
    bool RefinementOverride(double x, double p, int depth, int particleType) const {
        return hooks.RefinementOverride && hooks.RefinementOverride(x, p, depth, particleType);
    }
    // the inputs as the C description (what checkpoints record)
    vx_settings_desc Describe() const;
    // Settings.cpp:60-125: the run banner (configuration, particles, output
    // switches), in the reference's layout
    std::string Banner() const;

private:
    std::shared_ptr<const vx::Scenario> scenario_;  // keeps a scenario's hooks' target alive
    void finish();
};

// Shared device context (one GPU, one stream)
struct Device {
    vx_ctx *ctx = nullptr;
    explicit Device(int device);
    ~Device();
    void check(int rc, const char *what) const;
};

// Rectangle.hpp:6-80: one patch whose state lives in HBM.
class Rectangle {
public:
    int n_x, n_p, x_pos, p_pos;
    double relativeToBottom;
    int depth, particleType;
    bool up, down, left, right;
    // adaptive-mesh links (Rectangle.hpp:7-11); a null neighbour is the
    // domain-edge sentinel (BoundaryCondition)
    std::vector<Rectangle *> ext_xm, ext_xp, ext_pm, ext_pp;
    std::vector<char> same_xm, same_xp, same_pm, same_pp;
    std::vector<Rectangle *> finer_level, finer_level_x, finer_level_p;
    std::vector<char> is_nested, is_ilb_x, is_ilb_p, is_interpolated;
    double errorWeights[5];
    int table_index = -1;  // index in the mesh's device patch table
    int NS(int i, int j) const { return (n_p + 4) * (i + 2) + 2 + j; }
    void CalculateConnectivitySame(Rectangle *rect);                    // Rectangle.cpp:671-766
    void CalculateConnectivityFromFiner(Rectangle *current, Rectangle *rect);  // 768-864
    void getError(std::vector<std::pair<int, int>> &flagged);          // 866-890
    // flagged cells per x row / p column of the last getError's range (from the
    // device compaction), the range starting at (sig_lo_x, sig_lo_p)
    std::vector<int> sig_x, sig_p;
    int sig_lo_x = 0, sig_lo_p = 0;
    void UploadNested();
    Rectangle(int n_x, int n_p, int x_pos, int p_pos, int depth, Settings &settings,
              std::shared_ptr<Device> dev, bool up, bool down, bool left, bool right,
              int particleType);
    ~Rectangle();
    void SetFieldSolver(EMFieldSolver *solver) { EMSolver = solver; }
    void InitializeDistribution();
    void FCTTimeStep(double timestep, int step, int subStep);
    void FCTStageWalls(double timestep, int step);  // sub-step kernels, whole stage (walls only)
    void FCTStageFused(double timestep, int step, bool mixed = false);  // one fused kernel per stage (walls only)
    // the same stage as an item of a batched launch (vx_fct_stage_fused_batch);
    // the planes rotate as FCTStageFused rotates them
    void FusedItem(double timestep, int step, bool mixed, vx_stage_item *item);
    void UnaliasSlots();  // give f[0] and f[1] distinct planes (before sub-step kernels)
    int slot_version = 0;  // bumped whenever the f plane pointers change
    // the f[0..2] plane pointers and their version (saved and restored around
    // the recording of a step graph, which moves them without running)
    struct Slots {
        double *f[3];
        int version;
    };
    Slots GetSlots() const { return {{P.f[0], P.f[1], P.f[2]}, slot_version}; }
    void SetSlots(const Slots &s) {
        for (int k = 0; k < 3; k++) P.f[k] = s.f[k];
        slot_version = s.version;
    }
    void CalculateRhoAndJ();
    bool AllWalls() const { return up && down && left && right; }
    const vx_patch &desc() const { return P; }
    std::vector<double> Download(const double *plane, int ghosts) const;
    std::vector<double> DownloadSlot(int slot, int ghosts) const;  // vx_patch_download
    const double *Plane(const std::string &name) const;
    void UploadSlot(int slot, const std::vector<double> &full);  // whole plane incl. ghosts
    void UploadSlot(int slot, int ghosts, const double *host);  // (n_x+2g) x (n_p+2g), vx_patch_upload
    std::vector<double> CalculateEnergy();  // Rectangle.cpp:284-305 on the device
    double *energyR_d = nullptr;
    double fScale = 0.0;  // max |f| of the host uploads (scale of the mixed-precision FP32 copies)

private:
    Settings &settings;
    std::shared_ptr<Device> dev;
    EMFieldSolver *EMSolver = nullptr;
    vx_patch P{};
    uint8_t *nested_d = nullptr;
    double *buf[3] = {nullptr, nullptr, nullptr};  // physical planes behind f[0..2]
    std::vector<void *> allocations;
    double *alloc_plane();
};

// BoundaryCondition.hpp:7-11: the domain-edge neighbour whose same-level value
// is always 0.  On the device it is the zero ghost fill (vx_ghost_zero).
class BoundaryCondition {
public:
    double GetValueFromSameLevel(int, int, int = 1) const { return 0.0; }
};

// Level.hpp:4-23
class Level {
public:
    std::vector<std::shared_ptr<Rectangle>> rectangles;
    Level(int particleType, int depth, Settings &settings);
    ~Level();
    void FCTTimeStep(double timestep, int step, int subStep);
    // Level.cpp:88-126: 0 restriction, 1 same-level ghosts, 2 coarse-fine
    // ghosts, 3 corners, 4-6 limiter-C sync; each a device task program
    void PushData(int updateType, int val = 1);
    void RunFluxProgram(int step, double dt, const vx_fields &fl);
    int depth;
    // set by the mesh when the hierarchy changes
    Device *dev = nullptr;
    vx_patch_table *table = nullptr;
    vx_program *prog[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool prog_built[7] = {false, false, false, false, false, false, false};  // built on first use
    vx_program *flux_prog = nullptr;  // coarse-fine faces of this level's patches
    void FreePrograms();

private:
    int particleType;
    Settings &settings;
};

typedef std::pair<int, int> coords;  // veritas.hpp:45-48
typedef std::pair<coords, coords> rect;
typedef std::vector<rect> level;
// Mesh::splitRectangle (Mesh.cpp:644-792): the boxes that cover the flagged
// cells of box b (host/amr.cpp)
level SplitBox(const rect &b, const std::vector<coords> &flagged, double minEfficiency);
// Mesh.cpp:315-339 for one box (host/amr.cpp)
void cells_under_box(int nmax, int pmax, double fx, double fp, int i0, int i1, int j0, int j1,
                     std::vector<coords> &out);

// Mesh.hpp:4-41
class Mesh {
public:
    int particleType;
    std::vector<std::unique_ptr<Level>> levels;  // levels[0] is the finest
    std::vector<level> hierarchy;                // hierarchy[0] is the coarsest
    Mesh(int particleType, Settings &settings, std::shared_ptr<Device> dev);
    ~Mesh();
    void SetFieldSolver(EMFieldSolver *solver);
    void Advance(double timeStep, int step);
    void PushData(int val = 1);
    void PushBoundaryC();
    void updateHierarchy(bool init = false);       // Mesh.cpp:132-210
    void getError(int lvl, bool init, std::vector<coords> &flagged);  // Mesh.cpp:212-295
    bool SinglePatch() const;
    void Restore(const std::vector<level> &h);  // rebuild patches for a saved hierarchy
    // Mesh.cpp:877-902: f slot 0 of every patch; returns the formatted files
    std::vector<std::pair<std::string, std::string>> rectangleData(double tidx) const;
    void Sub0(double timeStep, int step);  // sub-step 0 of every patch with the coarse-fine fluxes
    int regrids = 0, rebuilds = 0;  // updateHierarchy calls after set-up, and those that changed the patches
    bool fused = true;  // single walled patch: one fused call per stage
    // false: every sub-step launched patch by patch (the plain per-patch form of
    // Level::FCTTimeStep, Level.cpp:12-17) instead of one launch per table
    bool batched = true;
    // mixed precision (FP32 WENO weights and Zalesak ratios): the fused stage of
    // a single walled patch, else the batched flux and limiter launches
    bool mixed = false;
    double FScale() const;  // scale of the FP32 copies: max over patches of max |f|
    // task batches (launches) of the mesh's programs and the tasks in each:
    // PushData's, PushBoundaryC's, then the coarse-fine flux program of each level
    std::vector<std::vector<int32_t>> ProgramBatches() const;

private:
    Settings &settings;
    std::shared_ptr<Device> dev;
    EMFieldSolver *EMSolver = nullptr;
    // device patch table and task programs of the current hierarchy (host/amr.cpp)
    vx_patch_table *table = nullptr;
    std::vector<Rectangle *> order;  // patches in table order
    // PushData's and PushBoundaryC's sequences of level programs, each merged
    // into one program (batches span levels and types where no cell conflicts)
    vx_program *push_prog = nullptr, *pushc_prog = nullptr;
    // pushc_prog followed by one batch of VX_OP_CORRECT tasks for every cell
    // next to a face whose C it writes (after vx_fct_sub12_table)
    vx_program *pushcc_prog = nullptr;
    void promoteHierarchyToMesh(bool init);      // Mesh.cpp:794-875
    void InterMeshDataTransfer(std::vector<std::unique_ptr<Level>> &levels_n);  // Mesh.cpp:116-130
    void BuildPrograms();
    void FreePrograms();
    void RefreshTable();
    void EnsureTable();  // refresh the device table if plane pointers moved
    int table_version = -1;
    friend class SolverManager;  // saves and restores table_version around step graphs
};

// EMSolver.hpp:9-49
class EMFieldSolver {
public:
    EMFieldSolver(Settings &settings, const std::vector<std::shared_ptr<Mesh>> &meshes,
                  std::shared_ptr<Device> dev);
    ~EMFieldSolver();
    void AssembleRhoAndJ();
    void UpdatePotential();
    void RGKStep(int step, double timestep);
    void EnforceChargeNeutralization();
    double EstimateCFLBound();
    const vx_fields &fields() const { return fl; }
    const vx_field_state &state() const { return fs; }
    void RebuildMomentPlan();
    void SnapshotStage0Fields();  // E and A^2 of RK stage 0 for the fused stage
    // output (host/output.cpp)
    void AssembleEnergy();                          // EMSolver.cpp:124-131
    std::vector<std::vector<double>> energies;      // dNdP per species, finest p grid
    std::vector<double> Download(const double *d, size_t n) const;
    double *charges_species = nullptr;              // n_species * n_f, last assembly
    // step graphs: RGKStep reads the hook values of stage s from drive_d[2s..2s+1]
    // (device memory written before each replay) instead of taking them as arguments
    bool device_drive = false;
    double *drive_d = nullptr;
    bool compensated_moments = false;  // mixed-precision form: Neumaier p sums

private:
    Settings &settings;
    std::vector<std::shared_ptr<Mesh>> meshes;
    std::shared_ptr<Device> dev;
    vx_field_state fs{};
    vx_fields fl{};
    vx_moment_plan *plan = nullptr;
    std::vector<void *> allocations;
    double *alloc(size_t n);
};

// SolverManager.hpp:6-19
class SolverManager {
public:
    explicit SolverManager(Settings &settings, int device = 0);
    void Advance(double timeStep);
    void AdvanceFields(double timeStep);
    double CalculateDt(double cfl) { return cfl * EMSolver->EstimateCFLBound(); }
    // SolverManager.cpp:48-58: regrid every species and add a row (t in
    // laser periods tempEM[0] / c, as the reference prints) to the screen table
    void reGrid(double t);
    // SolverManager.cpp:9-22 and 73-100: the progress table the reference
    // prints (header, the set-up row, a row per regrid); collected here, and
    // echoed to stdout when echo_screen is set
    const std::string &Screen() const { return screen_; }
    bool echo_screen = false;
    // SolverManager.cpp:103-160 and 64-70: the arrays are copied from the
    // device, then formatted and appended to output/ on a writer thread
    void fileOutput(double t);
    void OutputRectangles(double t);
    void FlushOutput();  // wait for the writer thread
    // checkpoint/restart (host/checkpoint.cpp; the reference has none): the
    // hierarchy, f slots 0 and 1 of every patch, the field state and the clock
    void SaveCheckpoint(const std::string &path);
    void LoadCheckpoint(const std::string &path);
    ~SolverManager();
    // pieces of Advance for stage-by-stage checks
    void StageMoments() { EMSolver->AssembleRhoAndJ(); }
    void StagePoisson() { EMSolver->UpdatePotential(); }
    void StageVlasov(double dt, int step);
    void StageVlasovSub(double dt, int step, int sub);
    void StageFields(double dt, int step);
    void SetFused(bool on);
    void SetBatched(bool on);
    // One step of Advance in pieces, for a caller that interleaves several runs
    // (the serving loop, host/serve.cpp): per stage StageBegin, then either the
    // batched item (FusedItem true) or StageVlasov, then StageEnd.
    void StageBegin(double dt, int step);  // moments, Poisson, stage-0 snapshot
    bool FusedItem(double dt, int step, vx_stage_item *item);
    void StageEnd(double dt, int step);    // stage clock and the Maxwell stage
    void CflAsync();                       // vx_cfl_bound_async for the next dt
    double CflResult(double cfl);          // cfl * the bound (CalculateDt)
    bool CflReady();                       // the pending bound has arrived
    // mixed-precision form: FP32 WENO weights and Zalesak ratios (the fused
    // stage of a single walled patch, the batched flux and limiter launches of
    // a hierarchy), compensated FP64 moment sums; FP64 otherwise
    void SetMixed(bool on);
    // Advance as a graph launch (host/graph.cpp): the first step from a given
    // plane-pointer state and dt runs eagerly, the second records the step's
    // launches once, later ones replay the recording.  On by default.
    void SetGraphs(bool on);
    int graph_replays = 0;   // steps run from a recorded graph
    int graph_records = 0;   // graphs recorded
    std::shared_ptr<Device> dev;
    std::vector<std::shared_ptr<Mesh>> meshes;
    std::shared_ptr<EMFieldSolver> EMSolver;
    Settings &settings;

private:
    std::string screen_;
    void ScreenWrite(const std::string &s);
    void ScreenRow(const Mesh &mesh);  // SolverManager.cpp:73-100
    std::shared_ptr<struct OutputFiles> files;
    std::thread writer;
    void StepKernels(double timeStep);  // SolverManager.cpp:28-39, launched eagerly
    // host state that a step moves: plane pointers, table versions, the clock
    struct HostState {
        std::vector<Rectangle::Slots> slots;
        std::vector<int> table_versions;
        double time = 0.0;
    };
    struct StepGraph {
        std::vector<uintptr_t> key;
        HostState post;  // after the step (the clock is advanced by the replay itself)
        vx_graph *g = nullptr;
        ~StepGraph() { vx_graph_destroy(g); }
    };
    std::vector<std::unique_ptr<StepGraph>> step_graphs;
    std::vector<std::vector<uintptr_t>> graph_seen, graph_refused;
    bool graphs_on = true;
    std::vector<uintptr_t> GraphKey(double dt) const;
    HostState SaveHostState() const;
    void RestoreHostState(const HostState &h);
    StepGraph *RecordStep(double dt, const std::vector<uintptr_t> &key);
    void DropGraphs();
};

}  // namespace vxh
