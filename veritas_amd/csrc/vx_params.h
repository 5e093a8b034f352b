/* Plain-C description of a run, shared by the C ABI (include/veritas_amd.h),
 * the Python bindings (ctypes) and the CPU oracle.  Field meanings follow the
 * reference's Input/Particles structs (Settings.hpp:4-31) plus the scenario
 * parameters that the reference leaves to user hooks (veritas.cpp:7-115).
 * All physical values are normalised (c/omega_p, m_e c, m_e c^2); the library
 * converts them to SI with the reference's constants. */
#ifndef VX_PARAMS_H
#define VX_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vx_scenario_params {
    int32_t kind;             /* 0 Landau, 1 two-stream, 2 laser */
    int32_t nx, np;           /* coarsest-level cells in x and p */
    int32_t levels;           /* Input::Lfinest */
    int32_t r;                /* Input::r, refinement ratio */
    int32_t quadrature_depth; /* Settings::quadratureDepth */
    int32_t modes;            /* two-stream perturbation modes 1..modes */
    int32_t pad0;
    uint64_t seed;            /* two-stream phases */
    double cfl;
    double min_efficiency, refinement_criteria, size_weight;
    double pre_length, post_length; /* c/omega_p */
    double n_ref;             /* m^-3 */
    double length;            /* domain length, c/omega_p */
    double pmin, pmax;        /* m_e c */
    double theta;             /* m_e c^2 */
    double amplitude, k, drift;
    double a0, n_over_nc, ramp, fwhm_cycles, lambda0;
    double gap;               /* laser: vacuum gap (wavelengths) splitting the plateau, 0 = none */
} vx_scenario_params;

/* The user hooks of the reference (Settings.hpp:45-64, written by the user in
 * veritas.cpp:36-115) as C function pointers; each receives `user` first.
 * Positions, momenta and times are SI.  A NULL get_by/get_bz drives nothing
 * (0), a NULL refinement_override never flags, a NULL settings_override
 * changes nothing; initial_distribution is required. */
typedef struct vx_settings_desc vx_settings_desc;
typedef struct vx_hooks {
    void *user;
    /* Settings::InitialDistribution(x, p, particleType) */
    double (*initial_distribution)(void *user, double x, double p, int32_t species);
    /* Settings::GetBY / GetBZ(x, t): transverse B at the left edge */
    double (*get_by)(void *user, double x, double t);
    double (*get_bz)(void *user, double x, double t);
    /* Settings::RefinementOverride(x, p, depth, particleType): nonzero flags the cell */
    int32_t (*refinement_override)(void *user, double x, double p, int32_t depth, int32_t species);
    /* Settings::settingsOverride(): may change any value of the description
     * after the inputs are copied and before they are checked (Settings.cpp:31) */
    void (*settings_override)(void *user, vx_settings_desc *settings);
} vx_hooks;

#define VX_MAX_SPECIES 8

/* Input and Particles of Settings.hpp:4-31 (SI), plus Settings::quadratureDepth
 * (set by settingsOverride in the reference, veritas.cpp:51). */
struct vx_settings_desc {
    double minEfficiency, dx, refinementCriteria, cfl, sizeWeight, preLength, postLength;
    uint32_t nx, r, Lfinest;        /* Input::nx, r, Lfinest */
    int32_t quadratureDepth;
    int32_t n_species;              /* entries used in the per-species arrays */
    int32_t pad0;
    double mass[VX_MAX_SPECIES], charge[VX_MAX_SPECIES];  /* Particles::mass, charge */
    double dp[VX_MAX_SPECIES], pmin[VX_MAX_SPECIES];      /* Particles::dp, pmin */
    uint32_t np[VX_MAX_SPECIES];                          /* Particles::np */
};

#ifdef __cplusplus
}
#endif

#endif
