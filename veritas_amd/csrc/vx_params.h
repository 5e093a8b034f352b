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

#ifdef __cplusplus
}
#endif

#endif
