/* veritas_amd: MI355X (gfx950) kernels for the Veritas Vlasov-Maxwell time
 * advance, behind a plain C ABI.
 *
 * The reference (Libbum/Veritas) has no FFI: its boundary is the C++ class API
 * SolverManager / Mesh / Level / Rectangle / EMFieldSolver (SolverManager.hpp,
 * Mesh.hpp, Level.hpp, Rectangle.hpp, EMSolver.hpp).  The C++ host library in
 * veritas_amd/csrc/host keeps those classes and calls the entry points below,
 * each of which replaces one reference routine (cited per function).
 *
 * Conventions: every call is stream-ordered on the context's HIP stream and
 * returns 0 on success or a negative VX_E* code; vx_last_error() gives the
 * message.  Pointers inside the descriptors are device pointers owned by the
 * caller (allocate them with vx_malloc).  Planes are FP64, laid out row-major
 * over x with p contiguous: element (i, j), i in [-2, n_x + 2), j in
 * [-2, n_p + 2), lives at (i + 2) * pitch + (j + 2); pitch is a multiple of 8.
 */
#ifndef VERITAS_AMD_H
#define VERITAS_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "../veritas_amd/csrc/vx_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VX_OK 0
#define VX_EINVAL -1   /* bad argument or shape */
#define VX_EHIP -2     /* HIP runtime error */
#define VX_ENOMEM -3   /* device allocation failed */
#define VX_ENOTSUP -4  /* configuration outside what the kernels handle */

typedef struct vx_ctx vx_ctx;

/* ---- context (no reference counterpart: the reference runs in-process) ---- */
int vx_ctx_create(int device, vx_ctx **out);
void vx_ctx_destroy(vx_ctx *ctx);
const char *vx_last_error(vx_ctx *ctx);
int vx_sync(vx_ctx *ctx);
void *vx_stream(vx_ctx *ctx); /* the hipStream_t all calls are ordered on */
int vx_malloc(vx_ctx *ctx, size_t bytes, void **dptr); /* zero-filled */
int vx_free(vx_ctx *ctx, void *dptr);
int vx_memcpy_h2d(vx_ctx *ctx, void *dst, const void *src, size_t bytes);
int vx_memcpy_d2h(vx_ctx *ctx, void *dst, const void *src, size_t bytes);
int vx_memcpy_d2d(vx_ctx *ctx, void *dst, const void *src, size_t bytes);
int vx_fill(vx_ctx *ctx, double *dst, double value, size_t count);

/* ---- one AMR patch (Rectangle, Rectangle.hpp:6-80) on the device ---- */
typedef struct vx_patch {
    int32_t n_x, n_p, x_pos, p_pos; /* Rectangle::n_x, n_p, x_pos, p_pos */
    int32_t rtb;                    /* Rectangle::relativeToBottom */
    int32_t pitch;                  /* doubles per x-row of every plane */
    int32_t left, right, down, up;  /* patch touches the domain edge (Rectangle.cpp:1257-1260) */
    double dx, dp, pmin, q, m;      /* cell sizes, Settings::pmin, charge, mass (SI) */
    /* derived exactly as the reference evaluates them (Rectangle.cpp:1271-1275,
     * Rectangle.hpp:181-183): 1/dx, 1/dp, cs*cs*m, (m_inv*c_inv)^2 */
    double dx_inv, dp_inv, cc, mc2inv;
    double *f[3];           /* Rectangle::f slots: 0 = f^n, 1 = current stage, 2 = LO predictor */
    double *FxH[6], *FpH[6]; /* high-order fluxes of each RK stage (Rectangle::FxH/FpH) */
    double *FxL0, *FpL0;    /* low-order fluxes of stage 0 (the only ones used, F8) */
    double *FxDS, *FpDS;    /* antidiffusive fluxes of the current stage */
    double *Rp, *Rm, *Cx, *Cp; /* Zalesak limiter state */
    uint8_t *nested;        /* per cell (plane layout, 1 byte): covered by a finer patch */
    double *chargeR, *currentR; /* n_x * rtb each */
} vx_patch;

/* Rectangle::f slot `slot` (0: f^n, 1: current stage, 2: predictor;
 * Rectangle.hpp:91-108 interleaves them, here each is a plane) of one patch
 * to or from a host array of (n_x + 2g) x (n_p + 2g) doubles, row-major, p
 * fastest, keeping g = 0, 1 or 2 ghost layers; waits for the copy.  For
 * output, initial data and golden files. */
int vx_patch_upload(vx_ctx *ctx, const vx_patch *p, int slot, int ghosts, const double *host);
int vx_patch_download(vx_ctx *ctx, const vx_patch *p, int slot, int ghosts, double *host);

/* 1-D finest-grid field data the patches read during a stage. */
typedef struct vx_fields {
    int32_t n_f;       /* Settings::x_size_finest */
    int32_t n_prepad;  /* EMFieldSolver::n_prepad */
    int32_t n_e;       /* n_f + n_prepad + n_postpad */
    int32_t pad;
    const double *E;   /* EMFieldSolver::GetEfield(i), i in [0, n_f) (periodic) */
    const double *a2;  /* EMFieldSolver::a_squared, n_f + 1 faces */
    const double *Ay1, *Az1; /* transverse potentials, RK slot 1, n_e cells */
    const double *E0, *a20;  /* E and a_squared as they were at RK stage 0 of this
                                step (the fused stage recomputes the stage-0
                                low-order fluxes from them) */
} vx_fields;

/* Plane geometry helper: pitch (doubles) for a patch with n_p cells. */
int32_t vx_plane_pitch(int32_t n_p);
size_t vx_plane_elems(int32_t n_x, int32_t n_p);

/* FCT sub-step 0 (Rectangle::FCTTimeStep subStep 0, Rectangle.cpp:1269-1535):
 * WENO face values, high- and low-order fluxes of RK stage `step`, the RK flux
 * combination (FxDS/FpDS) and the low-order predictor f[2]. */
int vx_fct_sub0(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl);
/* FCT sub-step 1: Zalesak R+-, then C on interior faces (Rectangle.cpp:1536-1594). */
int vx_fct_sub1(vx_ctx *ctx, const vx_patch *p);
/* FCT sub-step 2: f[1] = f[2] + limited antidiffusive fluxes (Rectangle.cpp:1595-1613). */
int vx_fct_sub2(vx_ctx *ctx, const vx_patch *p);
/* FCT sub-step 3: f[0] <- f[1] including ghosts (Rectangle.cpp:1614-1622). */
int vx_fct_sub3(vx_ctx *ctx, const vx_patch *p);

/* Whole RK stage of one patch whose four edges are domain walls (a single
 * unrefined level): sub-steps 0-2 with the zero ghost fills in between, fused
 * (Mesh::Advance, Mesh.cpp:64-89, for one patch).  Equivalent to
 * sub0 + ghost(2) + sub1 + sub2 + ghost(1) [+ sub3 at step 5]. */
int vx_fct_stage_walls(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl);

/* The same stage as one fused kernel (register streaming along x, wavefront
 * shuffles along p); store-flux design: recomputes the stage-0 low-order
 * fluxes from f[0] and the stage-0 fields, stores FxH/FpH of this stage, and
 * writes f_{s+1} into `fout`.  `fin` is f_s (p->f[1]); `fout` must alias
 * neither `fin` nor p->f[0]; ghost rings of all three planes must be zero.
 * At step 4 the FxH/FpH planes of step 4 receive step 5's partial RK sum
 * (its coefficients times the same dt, over the fluxes of steps 0-4) instead
 * of the step-4 fluxes, and step 5 reads that sum: run steps 4 and 5 of a
 * time step both fused, with one dt.  Bit-identical to the sub-step kernels.
 * Call vx_fct_stage_fused_available() to check that the library provides it. */
int vx_fct_stage_fused(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl,
                       const double *fin, double *fout);
/* Mixed-precision form of vx_fct_stage_fused (no counterpart in the
 * reference): the WENO weights of both face reconstructions and the Zalesak
 * ratios are evaluated in FP32 on f scaled by the power of two nearest below
 * f_scale (pass the patch's max |f|); candidate face values,
 * fluxes, the RK combination and every update of f stay FP64.  Not
 * bit-identical to the FP64 path: its error against the FP64 oracle is
 * measured by the tests. */
int vx_fct_stage_fused_mp(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl,
                          const double *fin, double *fout, double f_scale);
/* One patch of a batched stage launch: a walled single patch of an
 * independent run, with its own fields, planes and dt. */
typedef struct vx_stage_item {
    vx_patch p;
    vx_fields fl;
    const double *fin; /* f_s */
    double *fout;      /* f_(s+1) */
    double dt;
    double f_scale;    /* > 0: the mixed-precision stage with this f scale; 0: FP64 */
} vx_stage_item;
/* vx_fct_stage_fused (or _mp) for up to 64 patches at the same RK stage in
 * one launch (two when FP64 and mixed items are mixed): each patch's results
 * are bit-identical to its own launch.  The item arguments are staged through
 * pinned memory on the context stream, so the patches' inputs must be ready
 * on that stream (vx_ctx_wait).  No counterpart in the reference: the
 * serving loop packs the patches of concurrent runs into shared launches. */
int vx_fct_stage_fused_batch(vx_ctx *ctx, const vx_stage_item *items, int n, int step);
/* Rows per strip that the fused stage uses at RK stage `step` for patches of
 * n_x x n_p cells, in FP64 (mixed = 0) or mixed precision (mixed = 1); timed
 * on the first call for that shape, stage and precision, 0 before it.  No
 * counterpart in the reference: it reports the tile shape the library chose. */
int vx_stage_tile_rows(vx_ctx *ctx, int32_t n_x, int32_t n_p, int step, int mixed);
/* Resident waves per SIMD of the kernel form chosen with those strips: 4 (the
 * form compiled for at most 128 VGPRs) or 3 (the other), 0 before the first
 * call; both forms give the same bits.  No counterpart in the reference. */
int vx_stage_tile_waves(vx_ctx *ctx, int32_t n_x, int32_t n_p, int step, int mixed);
int vx_fct_stage_fused_available(void);

/* Ghost fill of slot `slot` of a patch whose neighbours are all the domain-edge
 * sentinel (BoundaryCondition::GetValueFromSameLevel returns 0,
 * BoundaryCondition.cpp:6-8): Rectangle::UpdateSameLevelBoundaries +
 * UpdateCornerPoints (Rectangle.cpp:562-614, 1130-1214) for that case. */
int vx_ghost_zero(vx_ctx *ctx, const vx_patch *p, int slot);

/* p-moments of one patch (Rectangle::CalculateRhoAndJ, Rectangle.cpp:157-282):
 * chargeR and currentR, summed over p in the reference's order. */
int vx_moments_patch(vx_ctx *ctx, const vx_patch *p, const vx_fields *fl);
/* The same moments with each p sum compensated: every lane accumulates its
 * p-cells with error-free additions (TwoSum), then one butterfly of TwoSums
 * combines the lanes' sums and errors (the
 * mixed-precision form keeps the moment reductions FP64 and compensated; the
 * reference's sequential order is not kept; no counterpart in the reference). */
int vx_moments_patch_compensated(vx_ctx *ctx, const vx_patch *p, const vx_fields *fl);
/* Rectangle::CalculateEnergy (Rectangle.cpp:284-305): energyR[j * rtb + k]
 * (n_p * rtb doubles, device) = dx * sum over non-nested cells (i, j) of the
 * p-sub-cell value k, summed in increasing i (the reference's loop races on
 * energyR; this is its sequential order). */
int vx_energy_patch(vx_ctx *ctx, const vx_patch *p, double *energyR);
/* Level/Mesh/EMFieldSolver assembly of per-patch moments onto the finest grid
 * (Level.cpp:42-62, Mesh.cpp:52-56, EMSolver.cpp:104-122), patches given in the
 * reference's order (species, then levels finest->coarsest, then rectangles).
 * level_start[s * n_levels + l] indexes the first patch of level l of species s
 * in `patches`.  A plan holds the device copy of that table; build it after
 * every change of the hierarchy. */
typedef struct vx_moment_layout {
    int32_t n_species;
    int32_t n_levels;            /* per species */
    const int32_t *level_start;  /* host array, n_species * n_levels + 1 entries */
} vx_moment_layout;
typedef struct vx_moment_plan vx_moment_plan;
int vx_moment_plan_create(vx_ctx *ctx, const vx_patch *patches, const vx_moment_layout *lay,
                          int32_t n_f, vx_moment_plan **out);
void vx_moment_plan_destroy(vx_moment_plan *plan);
/* charge, J: n_f each; charges_species (optional): n_species * n_f */
int vx_moments_assemble(vx_ctx *ctx, const vx_moment_plan *plan, double *charge, double *J,
                        double *charges_species);

/* ---- 1-D field solver (EMFieldSolver, EMSolver.hpp:9-49) ---- */
typedef struct vx_field_state {
    int32_t n_f, n_prepad, n_postpad, n_e;
    double dx_f;      /* finest dx */
    double dx2;       /* pow(dx_f, 2.0) as the host evaluates it (EMSolver.cpp:163) */
    double field_coef; /* 1 / (12 dx_f) (EMSolver.cpp:88) */
    double *charge, *J, *neutral; /* n_f each */
    double *phi;      /* n_f */
    double *E;        /* n_f: GetEfield(i) after the last Poisson solve */
    double *Ex0;      /* 1 scalar on the device */
    double *By, *Bz, *Ey, *Ez, *Ay, *Az; /* 8 RK slots * n_e each, slot-major */
    double *a2;       /* n_f + 1 */
    double *work;     /* scratch, vx_field_work_elems(n_f) doubles */
} vx_field_state;
size_t vx_field_work_elems(int32_t n_f);

/* EMFieldSolver::UpdatePotential (EMSolver.cpp:156-192): right-hand side,
 * Poisson solve of the reference's matrix (EMSolver.cpp:28-88) to near correct
 * rounding, E on the finest grid and the Ex0 update. */
int vx_poisson(vx_ctx *ctx, const vx_field_state *fs);
/* vx_poisson that allocates nothing and needs no workspace: the same kernels
 * and bits, with the multi-workgroup solve (n_f > 4096) working in a fixed
 * device array of the library instead of fs->work (which may be NULL).  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code recorded into a graph, tight serving loops, harnesses that
 * forbid hipMalloc).  n_f above VX_POISSON_NOALLOC_MAX_NF is rejected
 * (VX_EINVAL).  The array (3.67 MB) is held by each GPU the library runs on
 * until the process ends and is shared by every call on that GPU: calls that
 * overlap on one GPU (other contexts, streams or threads) race.  Nothing in it
 * needs to start at zero: every call writes each entry before reading it. */
#define VX_POISSON_NOALLOC_MAX_NF 65536
int vx_poisson_noalloc(vx_ctx *ctx, const vx_field_state *fs);
/* EMFieldSolver::RGKStep (EMSolver.cpp:194-202): transverse-field RHS, RK
 * update and WENO A^2 at faces.  bz0/by0 are the GetBZ/GetBY hook values at
 * x = 0 for this stage's time (veritas.cpp:80-104). */
int vx_maxwell_stage(vx_ctx *ctx, const vx_field_state *fs, int step, double dt, double by0,
                     double bz0);
/* The same stage with the hook values read from device memory when the
 * kernels run (drive[0] = by0, drive[1] = bz0), so that a recorded step graph
 * takes new hook values per replay (EMSolver.cpp:492-499 evaluates the hooks
 * at every stage). */
int vx_maxwell_stage_dev(vx_ctx *ctx, const vx_field_state *fs, int step, double dt,
                         const double *drive);
/* EMFieldSolver::EnforceChargeNeutralization tail (EMSolver.cpp:624-628):
 * J = 0, neutral = -charge. */
int vx_neutralize(vx_ctx *ctx, const vx_field_state *fs);
/* EMFieldSolver::EstimateCFLBound (EMSolver.cpp:631-664) for n_species species
 * (host arrays q, m, dp_finest); result written to *out on the host. */
int vx_cfl_bound(vx_ctx *ctx, const vx_field_state *fs, int n_species, const double *q,
                 const double *m, const double *dp_finest, double *out);
/* The same bound computed on the device and copied to pinned host memory
 * without waiting; vx_cfl_result waits for that copy and returns the bound
 * (bit-identical to vx_cfl_bound: the maximum is exact in any order).  Lets a
 * caller that steps several runs queue all of them before it waits for one. */
int vx_cfl_bound_async(vx_ctx *ctx, const vx_field_state *fs, int n_species, const double *q,
                       const double *m, const double *dp_finest);
int vx_cfl_result(vx_ctx *ctx, double *out);
/* *ready = 1 when the pending bound has arrived (vx_cfl_result will not wait). */
int vx_cfl_ready(vx_ctx *ctx, int *ready);

/* ---- adaptive mesh: cell-task programs over a table of patches ----
 *
 * The reference fills ghost cells, restricts, synchronises the limiter C and
 * substitutes coarse-fine fluxes with per-rectangle loops over neighbour
 * pointers (Rectangle.cpp:314-614, 892-941, 1032-1253, 1625-1960; Level.cpp:
 * 88-150; Mesh.cpp:91-130, 904-917).  Here the host walks the same loops once
 * per hierarchy and records each cell or face operation as a task; a program
 * is a list of tasks cut into batches so that no task in a batch reads or
 * writes what another task of the same batch writes (the batches replay the
 * reference's sequential order exactly).  Each batch is one launch. */
typedef struct vx_patch_table vx_patch_table;
/* Device copy of n patch descriptors; tasks name patches by index. */
int vx_patch_table_create(vx_ctx *ctx, const vx_patch *patches, int32_t n, vx_patch_table **out);
/* Refresh the descriptors (same count) after plane pointers changed. */
int vx_patch_table_update(vx_ctx *ctx, vx_patch_table *t, const vx_patch *patches, int32_t n);
void vx_patch_table_destroy(vx_patch_table *t);

typedef struct vx_cell_task {
    int32_t op;        /* VX_OP_* */
    int32_t dst;       /* patch written */
    int32_t di, dj;    /* patch-local cell or face of dst */
    int32_t src;       /* patch read */
    int32_t si, sj;    /* op-specific source coordinates (below) */
    int32_t aux;       /* op-specific */
} vx_cell_task;

enum {
    /* f slot `slot` of dst at (di, dj) = ... (mode VX_MODE_GHOST), or slots 0
     * and 1 of dst = ... read from slot 1 of src (mode VX_MODE_TRANSFER) */
    VX_OP_ZERO = 0,     /* 0 (domain-edge sentinel, BoundaryCondition.cpp:6-8) */
    VX_OP_COPY = 1,     /* src at patch-local (si, sj) (GetValueFromSameLevel, Rectangle.cpp:307-312) */
    VX_OP_COARSE = 2,   /* element aux (= k * r + l) of src's conservative interpolation of
                           finest-level-of-src+1 cell (si, sj) (GetWenoValueFromCoarseLevel,
                           Rectangle.cpp:343-415) */
    VX_OP_RESTRICT = 3, /* mean of src's r x r cells under coarse cell (si, sj) given in the
                           coarse level's coordinates (GetValueFromFinerLevel, Rectangle.cpp:314-327) */
    /* limiter C (Rectangle.cpp:1625-1915); aux = t, the face type 0..3 */
    VX_OP_C_SAME = 4,   /* one face: dst = neighbour (di, dj), src = caller (si, sj) (SetCFromSameLevel) */
    VX_OP_C_DIFF = 5,   /* dst = coarse (di, dj) = (i_coarse, j_coarse), src = finer caller with
                           (si, sj) = (i_caller, j_caller) (SetCFromDifferentLevel) */
    VX_OP_C_UPDATE = 6, /* one face: min of both sides (UpdateCFromSameLevel) */
    /* coarse face (di, dj) of dst takes the finer patch src's fluxes over the r
     * faces starting at (si, sj) (CalculateFluxToCoarse*, Rectangle.hpp:132-178)
     * and its antidiffusive flux is recombined; bit k of aux set: finer face k
     * is itself a coarse-fine face, whose stored low-order flux is used */
    VX_OP_FLUX_X = 7,
    VX_OP_FLUX_P = 8,
    /* f slot 1 of dst at interior cell (di, dj) = its corrected value
     * (Rectangle.cpp:1596-1612) from f slot 2, C and the antidiffusive fluxes
     * of dst: the correction of a cell next to a face whose C the limiter-C
     * synchronisation wrote, after vx_fct_sub12_table (no counterpart in the
     * reference, whose corrector runs after the synchronisation) */
    VX_OP_CORRECT = 9
};
enum { VX_MODE_GHOST = 0, VX_MODE_TRANSFER = 1 };

typedef struct vx_program vx_program;
/* Upload tasks; batch_start[b] is the first task of batch b (n_batches + 1
 * entries, the last = n_tasks). */
int vx_program_create(vx_ctx *ctx, const vx_cell_task *tasks, int32_t n_tasks,
                      const int32_t *batch_start, int32_t n_batches, vx_program **out);
void vx_program_destroy(vx_program *prog);
/* Number of task batches of a program (each one launch) and, when n is large
 * enough, the tasks in each. */
int32_t vx_program_batch_sizes(const vx_program *prog, int32_t *sizes, int32_t n);
/* Ghost-fill / restriction / regrid-transfer / limiter-C programs.  r is the
 * refinement ratio (2 or 4). */
int vx_program_run(vx_ctx *ctx, const vx_patch_table *table, const vx_program *prog, int mode,
                   int slot, int r);

/* Regrid transfer in rectangles (Mesh.cpp:116-130 with Rectangle.cpp:892-941
 * and 1100-1128, whose per-cell loops these replace): slots 0 and 1 of dst on
 * a rectangle of cells, from slot 1 of src.
 *   VX_OP_COPY:   dst(di0 + a, dj0 + b) = src(si0 + a, sj0 + b), a < ni, b < nj
 *   VX_OP_COARSE: coarse cells (ci, cj) = (si0 + a, sj0 + b) of src in level
 *                 coordinates, a < ni, b < nj, each to its r x r finer cells
 *                 dst((ci - di0) r + k, (cj - dj0) r + l) = element l + r k of
 *                 the REF interpolation of that coarse cell (di0, dj0: dst's
 *                 position in src's level coordinates)
 * One call is one phase of the transfer (one launch over every range; the
 * ranges of a call must write disjoint cells and read no cell a range of the
 * call writes); it waits for the launch.  No counterpart in the reference. */
typedef struct vx_range_task {
    int32_t op, dst, src;
    int32_t di0, dj0, si0, sj0, ni, nj;
    int32_t pad;
} vx_range_task;
int vx_transfer_ranges(vx_ctx *ctx, const vx_patch_table *table, const vx_range_task *ranges, int32_t n,
                       int32_t r);
/* Coarse-fine flux program of RK stage `step` (run after vx_fct_sub0_fluxes
 * of every patch and before vx_fct_sub0_predictor), finest coarse level first.
 * At step 4 it also writes stage 5's partial RK sum at the faces whose fluxes
 * it replaces (see vx_fct_sub0_fluxes_table). */
int vx_program_run_fluxes(vx_ctx *ctx, const vx_patch_table *table, const vx_program *prog,
                          int step, double dt, int r, const vx_fields *fl);

/* Work-list forms of the sub-step kernels (no counterpart in the reference,
 * which loops Level::FCTTimeStep over its rectangles, Level.cpp:12-17): one
 * launch covers every patch of a table, each workgroup finding its patch in a
 * per-table list built on the first call (a table's shapes never change).
 * Each patch's results are bit-identical to its own launch: within a sub-step
 * the patches only write their own planes.  The table must hold the current
 * plane pointers (vx_patch_table_update). */
/* The sub-step-0 flux launches (these and vx_fct_sub0_fluxes) at step 4 also
 * write stage 5's partial RK sum (its coefficients times the same dt over the
 * fluxes of steps 0-4, left to right) into the FxH[5]/FpH[5] planes, and at
 * step 5 read it there in place of the five earlier flux planes: run steps 4
 * and 5 of a time step with one dt. */
int vx_fct_sub0_fluxes_table(vx_ctx *ctx, vx_patch_table *t, int step, double dt, const vx_fields *fl);
int vx_fct_sub0_predictor_table(vx_ctx *ctx, vx_patch_table *t, int step, double dt, int32_t n_f);
int vx_fct_sub1_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f);
int vx_fct_sub2_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f);
int vx_fct_sub3_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f);
/* Mixed-precision forms of vx_fct_sub0_fluxes_table and vx_fct_sub1_table
 * (no counterpart in the reference): the WENO weights of both face
 * reconstructions and the Zalesak ratios R+- in FP32 on f scaled by the power
 * of two nearest below f_scale (the mesh's max |f|), as vx_fct_stage_fused_mp
 * does for one walled patch; candidate face values, fluxes, the RK
 * combination, the limiter C and every update of f stay FP64.  The other
 * sub-steps and the task programs are shared with the FP64 path. */
int vx_fct_sub0_fluxes_table_mp(vx_ctx *ctx, vx_patch_table *t, int step, double dt, const vx_fields *fl,
                                double f_scale);
int vx_fct_sub1_table_mp(vx_ctx *ctx, vx_patch_table *t, int32_t n_f, double f_scale);
/* The limiter and the corrector of every patch in one launch (sub-steps 1
 * and 2, Rectangle.cpp:1536-1612; no counterpart in the reference, which runs
 * them as two loops with the limiter-C synchronisation of Mesh::PushBoundaryC,
 * Mesh.cpp:904-917, between them).  Writes R, C and f slot 1 of every interior
 * cell from C before the synchronisation; the caller then runs a
 * synchronisation program whose last batch recomputes, with VX_OP_CORRECT
 * tasks, f slot 1 of every cell next to a face whose C the synchronisation
 * writes.  The result equals vx_fct_sub1_table, the synchronisation, then
 * vx_fct_sub2_table. */
int vx_fct_sub12_table(vx_ctx *ctx, vx_patch_table *t, int32_t n_f);
int vx_fct_sub12_table_mp(vx_ctx *ctx, vx_patch_table *t, int32_t n_f, double f_scale);

/* Split form of vx_fct_sub0 for hierarchies: the fluxes and RK combination
 * (Rectangle.cpp:1269-1517), then the low-order predictor (1518-1534). */
int vx_fct_sub0_fluxes(vx_ctx *ctx, const vx_patch *p, int step, double dt, const vx_fields *fl);
int vx_fct_sub0_predictor(vx_ctx *ctx, const vx_patch *p, int step, double dt);

/* Refinement flags of one patch (Rectangle::getError, Rectangle.cpp:866-890,
 * ErrorEstimate, Rectangle.hpp:128-130): flags[(i - lo_x) * (hi_p - lo_p) +
 * (j - lo_p)] = 1 where the error estimate of patch-local cell (i, j) in
 * [lo_x, hi_x) x [lo_p, hi_p) exceeds `criteria` (uint8 device array). */
int vx_flag_cells(vx_ctx *ctx, const vx_patch *p, const double *weights5, double criteria,
                  int32_t lo_x, int32_t hi_x, int32_t lo_p, int32_t hi_p, uint8_t *flags);
/* The same flags, compacted on the device (device regrid assistance; no
 * counterpart in the reference): `cells` receives the flagged cells' indices
 * (i - lo_x) * (hi_p - lo_p) + (j - lo_p) in increasing order (the row-major
 * scan order of Mesh.cpp:212-295), `count` their number (at most max_cells,
 * else an error), and row_sig / col_sig the signatures of the range, the
 * number of flagged cells in each x row (hi_x - lo_x entries) and p column
 * (hi_p - lo_p entries) (the counts of Mesh::computeSignatures,
 * Mesh.cpp:386-447).  Host pointers; any of cells, row_sig, col_sig may be
 * NULL. */
int vx_flag_compact(vx_ctx *ctx, const vx_patch *p, const double *weights5, double criteria,
                    int32_t lo_x, int32_t hi_x, int32_t lo_p, int32_t hi_p, int32_t *cells,
                    int32_t max_cells, int32_t *count, int32_t *row_sig, int32_t *col_sig);

/* ---- timing helpers (HIP events on the context stream) ---- */
int vx_event_record(vx_ctx *ctx, int slot);
int vx_event_elapsed_ms(vx_ctx *ctx, int slot_a, int slot_b, float *ms);
/* Bracket every launch of the named kernel (e.g. "k_sub0") with events until
 * vx_timing_end, which synchronises and returns the summed device time. */
int vx_timing_begin(vx_ctx *ctx, const char *kernel);
int vx_timing_end(vx_ctx *ctx, double *total_ms, int *launches);

/* ---- step graphs (SolverManager::Advance, SolverManager.cpp:28-39, replayed
 * as one hipGraph launch per step) ----
 * vx_graph_begin starts recording the context stream: entry points called
 * until vx_graph_end are recorded, not run.  Entry points that synchronise,
 * allocate or copy to or from pageable host memory fail while recording (and
 * the recording is then discarded by vx_graph_end, which returns the error).
 * A graph holds the kernel arguments it was recorded with; the caller replays
 * it only while those buffers and values still apply. */
typedef struct vx_graph vx_graph;
int vx_graph_begin(vx_ctx *ctx);
int vx_graph_end(vx_ctx *ctx, vx_graph **out);
int vx_graph_launch(vx_ctx *ctx, const vx_graph *g);
void vx_graph_destroy(vx_graph *g);
/* Stream-ordered upload of at most 256 bytes that does not wait for the
 * device: the bytes are staged in a pinned ring slot (a slot is reused only
 * after its previous copy has completed). */
int vx_upload_small(vx_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Order: the waiter's later work starts after everything enqueued so far on
 * the signaller's stream (an event on the signaller, a wait on the waiter). */
int vx_ctx_wait(vx_ctx *waiter, vx_ctx *signaller);
/* Device bytes the context holds through vx_malloc (freed by vx_free). */
size_t vx_ctx_live_bytes(vx_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif
