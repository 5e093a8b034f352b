// CPU ORACLE -- test infrastructure only.
//
// A plain FP64 restatement of the reference's Vlasov-Maxwell time advance
// (Libbum/Veritas).  It is the checker for the GPU framework: only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and
// never as the thing that is measured or shipped.
//
// Every routine cites the reference file:line it restates.  Loop ranges,
// operation order and association follow the reference literally, because the
// GPU kernels are required to reproduce the same bits.
//
// Deliberate, documented differences from the reference:
//  * Poisson (EMSolver.cpp:28-88,156-192): the reference factors the dense
//    matrix with LAPACK dgetrf and solves with dgetrs.  The oracle solves the
//    same matrix to (near) correct rounding: an O(N) bordered band solve
//    followed by iterative refinement with a double-double residual.  Pinned in
//    the tests against numpy.linalg.solve (LAPACK gesv = getrf + getrs).
//  * Logarithms in the current moment use vx_log (within 1 ulp of the C
//    library) instead of MKL's vdLn (Rectangle.cpp:227-229).
//  * The limiter-C synchronisation against the domain-edge sentinel
//    (Rectangle.cpp:1835-1877 with BoundaryCondition as the neighbour) reads and
//    writes the sentinel's empty arrays in the reference (undefined behaviour);
//    the only entries of a real patch it could change are wall faces, which the
//    update loops never use (Rectangle.cpp:1257-1260), so the oracle skips it.
//  * Rectangle::CalculateEnergy (Rectangle.cpp:284-305) races on energyR; the
//    oracle computes the race-free (sequential) sum.
#pragma once

#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "../veritas_amd/csrc/scenario.h"

namespace orc {

// Rectangle.cpp:76-89: interpolation matrix of the sub-cell cubic
inline constexpr double kInterp[12] = {
    0.104166666666667, -0.708333333333334, 0.708333333333334, -0.104166666666667,
    0.117647058823529, 0.029411764705882,  0.029411764705882, 0.117647058823529,
    -0.083333333333333, 0.166666666666667, -0.166666666666667, 0.083333333333334};

double weno_nomax(double f1, double f2, double f3, double f4, bool right);
double face_weno(double f1, double f2, double f3, double f4);

struct Patch;
struct Field;
struct Mesh;

// Settings.hpp:4-65 / Settings.cpp:5-195
// The user hooks (Settings.hpp:45-64) come as the C function pointers of
// vx_hooks; settingsOverride (Settings.cpp:31) edits the C description.
struct Settings {
    double dx, time, minEfficiency, preLength, postLength, refinementCriteria, cfl, sizeWeight;
    std::vector<double> m, q, dp, fMax, pmin;
    unsigned x_size_finest, x_size, refinementRatio;
    int maxDepth, quadratureDepth;
    std::vector<unsigned> p_size, p_size_finest;
    vx_hooks hooks;

    Settings(const vx_settings_desc &d, const vx_hooks &h);
    double GetDp(int level, int sp) const;
    double GetDx(int level) const;
    int GetXSize(int level) const;
    int GetPSize(int level, int sp) const;
    void UpdateTime(int step, double dt);
    void DetermineMaximum();
    double InitialDistribution(double x, double p, int sp) const {
        return hooks.initial_distribution(hooks.user, x, p, sp);
    }
    double GetBY(double x, double t) const { return hooks.get_by ? hooks.get_by(hooks.user, x, t) : 0.0; }
    double GetBZ(double x, double t) const { return hooks.get_bz ? hooks.get_bz(hooks.user, x, t) : 0.0; }
    bool RefinementOverride(double x, double p, int depth, int sp) const {
        return hooks.refinement_override && hooks.refinement_override(hooks.user, x, p, depth, sp) != 0;
    }
};

// Rectangle.hpp:6-80
struct Patch {
    int n_x, n_p, x_pos, p_pos, depth, sp, r;
    double rtb;  // relativeToBottom
    double dp, dx, m_inv;
    bool up, down, left, right;
    Settings *settings;
    Field *em = nullptr;

    std::vector<double> f[3];
    std::vector<double> ex, fx, ep, fp;
    std::vector<double> FxL[6], FxH[6], FpL[6], FpH[6];
    std::vector<double> FxLS, FpLS, FxDS, FpDS, Rp, Rm, Cx, Cp;
    std::vector<double> chargeR, currentR, energyR;
    std::vector<double> interpREL, interpREF, errorWeights;
    std::vector<char> is_nested, is_interpolated;
    // neighbours across the patch edges (nullptr = domain-edge sentinel);
    // *_same flags say whether the neighbour is on the same level
    std::vector<Patch *> ext_xm, ext_xp, ext_pm, ext_pp;
    std::vector<char> same_xm, same_xp, same_pm, same_pp;
    // finer patch covering a cell / owning a coarse-fine face (Rectangle.hpp:8)
    std::vector<Patch *> finer_level, finer_level_x, finer_level_p;
    std::vector<char> is_ilb_x, is_ilb_p;  // is_interrior_level_boundary_x/p

    Patch(int n_x, int n_p, int x_pos, int p_pos, int depth, Settings &s, bool up, bool down,
          bool left, bool right, int sp);
    int NS(int i, int j) const { return (n_p + 4) * (i + 2) + 2 + j; }
    double Momentum(double i) const;
    double Gamma(double p, double a2) const;
    int GetFinestIndex(int i) const { return (int)(rtb * (i + x_pos)); }
    double GetEfield(int i) const;
    double GetValueFromSameLevel(int i, int j, int val) const;
    std::vector<double> GetInterpolantsREL(double f1, double f2, double f3, double f4,
                                           double f5) const;
    void InitializeDistribution();
    void CalculateRhoAndJ();
    void CalculateEnergy();
    void FCTTimeStep(double dt, int step, int sub);
    void UpdateSameLevelBoundaries(int val);
    void UpdateCornerPoints(int val);
    void UpdateDifferentLevelBoundaries(int val);
    // adaptive mesh (oracle_amr.cpp)
    std::vector<double> GetInterpolantsREF(double f1, double f2, double f3, double f4,
                                           double f5) const;
    double GetValueFromFinerLevel(int i, int j, int val) const;
    void UpdateInterriorPoints(int val);
    std::vector<double> GetWenoValueFromCoarseLevel(int i, int j, int d, int val) const;
    static bool OnLine(int x, int y1, int y2) { return ((x - y1) > -1) && ((y2 - x) > -1); }
    void CalculateConnectivitySame(Patch *rect);
    void CalculateConnectivityFromFiner(Patch *current, Patch *rect);
    double ErrorEstimate(int i, int j) const;
    void getError(std::vector<std::pair<int, int>> &flagged) const;
    void GetDataFromCoarseLevelRectangle(Patch *rect) const;
    void GetDataFromSameLevelRectangle(Patch *rect) const;
    void GetDataFromCoarseNewLevelRectangle(Patch *rect) const;
    double RGKGetFluxX(int i, int j) const;
    double RGKGetFluxP(int i, int j) const;
    double RGKGetFluxXL(int i, int j) const;
    double RGKGetFluxPL(int i, int j) const;
    double CalculateFluxToCoarseX(int i, int j) const;
    double CalculateFluxToCoarseP(int i, int j) const;
    double CalculateFluxToCoarseXL(int i, int j) const;
    double CalculateFluxToCoarsePL(int i, int j) const;
    void SetCFromDifferentLevel(int i, int j, Patch *rect, int t);
    void CalculateDifferentBoundaryC();
    void SetCFromSameLevel(int i, int j, Patch *rect, int t);
    void CalculateSameBoundaryC();
    void UpdateCFromSameLevel(int i, int j, Patch *rect, int t);
    void UpdateSameBoundaryC();
};

// EMSolver.hpp:9-49
struct Field {
    Settings &settings;
    int n_prepad, n_postpad;
    unsigned x_size;
    std::vector<Mesh *> meshes;
    std::vector<double> charge, J, By, Bz, Ey, Ez, Ay, Az, a_squared, neutralizationCharge, PHI;
    std::vector<std::vector<double>> charges, energies;
    double fieldCoef, Ex0;
    double w1, w2, w3;
    // bordered band factorisation of the Poisson matrix (rows/cols 1..N-1)
    std::vector<double> bandL1, bandL2, bandD, bandU1, bandU2, Binv_u;
    double schur;

    Field(Settings &s, std::vector<Mesh *> meshes);
    int Index(int i, int step) const { return step * (int)(x_size + n_prepad + n_postpad) + i; }
    void AssembleRhoAndJ();
    void AssembleEnergy();
    double GetASquared(int i) const;
    double GetEfield(int i) const;
    double GetCellAverageASquared(int i) const;
    void UpdatePotential();
    void PoissonSolve(const std::vector<double> &b, std::vector<double> &x) const;
    void PoissonApprox(const std::vector<double> &b, std::vector<double> &x) const;
    void PoissonResidual(const std::vector<double> &b, const std::vector<double> &x,
                         std::vector<double> &res) const;
    void RGKStep(int step, double dt);
    void RGKCalculateRHS(int step);
    void RGKUpdateIntermediateSolution(int step, double dt);
    void InterpolateToFaces();
    void EnforceChargeNeutralization();
    double EstimateCFLBound() const;
};

struct Level {  // Level.hpp:4-23
    int depth, sp;
    Settings &settings;
    std::vector<std::unique_ptr<Patch>> rectangles;
    std::vector<double> chargeL, currentL, energyL;
    Level(int sp, int depth, Settings &s);
    void FCTTimeStep(double dt, int step, int sub);
    void CollectRhoAndJ();
    void InterpolateRhoAndJToFinestMesh(std::vector<double> &charge, std::vector<double> &J);
    void CollectEnergy();
    void InterpolateEnergyToFinestMesh(std::vector<double> &energy);
    void PushData(int type, int val);
    void GetDataFromSameLevel(Level &lvl);
    void GetDataFromCoarserLevel(Level &lvl);
    void GetDataFromCoarseNewLevel(Level &lvl);
};

typedef std::pair<int, int> coords;        // veritas.hpp:45-48
typedef std::pair<coords, coords> rect;
typedef std::vector<rect> level;

struct Mesh {  // Mesh.hpp:4-41
    Settings &settings;
    int sp;
    Field *em = nullptr;
    std::vector<level> hierarchy;                // hierarchy[0] is the coarsest
    std::vector<std::unique_ptr<Level>> levels;  // levels[0] is the finest
    Mesh(int sp, Settings &s);
    void SetFieldSolver(Field *f);
    void Advance(double dt, int step);
    void PushData(int val = 1);
    void PushBoundaryC();
    void InterpolateRhoAndJToFinestMesh(std::vector<double> &charge, std::vector<double> &J);
    void InterpolateEnergyToFinestMesh(std::vector<double> &energy);
    // hierarchy construction and regrid (oracle_amr.cpp)
    void updateHierarchy(bool init = false);
    void getError(int lvl, bool init, std::vector<coords> &flagged);
    void interpRectanglesUp(level &identified, int lvl);
    void mergeDownFlaggedData(int lvl, const rect &r, std::vector<coords> &found);
    static int countCells(const rect &span);
    static void isFlaggedInside(const std::vector<coords> &flagged, std::vector<bool> &inside,
                                const rect &r);
    static std::tuple<std::vector<int>, std::vector<int>, std::vector<int>, std::vector<int>>
    computeSignatures(const rect &r, const std::vector<coords> &flagged);
    static coords identifyInflection(const rect &r,
                                     const std::tuple<std::vector<int>, std::vector<int>,
                                                      std::vector<int>, std::vector<int>> &sig);
    static std::tuple<bool, int, int> hasHole(const std::vector<int> &sig);
    static void getExtrema(rect &extrema, const std::vector<coords> &flagged);
    static level splitRectangle(const rect &r, const std::vector<coords> &flagged, double minEfficiency);
    void promoteHierarchyToMesh(bool init);
    void InterMeshDataTransfer(std::vector<std::unique_ptr<Level>> &levels_n);
};

struct Solver {  // SolverManager.hpp:6-19
    Settings settings;
    std::vector<std::unique_ptr<Mesh>> meshes;
    std::unique_ptr<Field> em;
    Solver(const vx_settings_desc &d, const vx_hooks &h);
    void Advance(double dt);
    void AdvanceFields(double dt);
    void reGrid();
    double CalculateDt(double cfl) { return cfl * em->EstimateCFLBound(); }
};

}  // namespace orc
