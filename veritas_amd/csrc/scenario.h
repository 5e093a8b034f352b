// Built-in scenarios: the inputs of a run (grid sizes, species, initial
// distribution, laser drive) for the three configurations benchmarked here.
// A scenario is one provider of what the reference leaves to the application
// (veritas.cpp:7-115: initialConditions, settingsOverride,
// InitialDistribution, GetBY, GetBZ, RefinementOverride): desc() gives the
// Input/Particles values and hooks() the hook functions, which any other
// caller supplies itself (vx_settings_desc, vx_hooks in vx_params.h).
//
// Everything here is an *input*: the GPU framework and the CPU oracle both read
// the same scenario so that they start from identical data.  Physical
// quantities are SI, using the reference's constants (veritas.hpp:22-29).
// Normalised configuration values (c/omega_p, m_e c) are converted with a
// reference density n_ref.
#pragma once

#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "vx_consts.h"
#include "vx_params.h"

namespace vx {

enum ScenarioKind { SCN_LANDAU = 0, SCN_TWO_STREAM = 1, SCN_LASER = 2 };

struct Scenario {
    int kind = SCN_LANDAU;
    // grid (coarsest level)
    int nx = 64, np = 256;
    int levels = 1;          // Lfinest
    int r = 2;               // refinement ratio
    int quadrature_depth = 1;
    double cfl = 0.4;
    double min_efficiency = 0.6, refinement_criteria = 1e-8, size_weight = 0.0;
    double pre_length = 0.0, post_length = 0.0;  // metres
    // normalised physics parameters
    double n_ref = 1.0e24;   // m^-3, defines omega_p and the unit of f
    double length = M_PI / 2;  // domain length in c/omega_p
    double pmin = -1.0, pmax = 1.0;  // m_e c
    double theta = 0.01;     // Maxwell-Juttner temperature (m_e c^2 units)
    double amplitude = 0.01; // density perturbation
    double k = 4.0;          // Landau wave number (omega_p/c units)
    double drift = 0.2;      // two-stream beam drift (m_e c)
    int modes = 8;           // two-stream perturbation modes 1..modes
    uint64_t seed = 42;
    // laser
    double a0 = 0.5, n_over_nc = 0.01, ramp = 50.0, fwhm_cycles = 10.0;
    double lambda0 = 1.0e-6;  // laser wavelength (m) for the laser scenario
    double gap = 0.0;         // vacuum gap (wavelengths) in the middle of the plateau
    // derived (filled by finalize)
    double omega_p = 0, unit_len = 0, unit_p = 0, norm = 0;
    double phases[64] = {0};
    double mass = m_electron, charge = -e_charge;

    // Exponentially scaled Bessel function K1e(z) = exp(z) K1(z), from
    // K1(z) = int_0^inf exp(-z cosh t) cosh t dt (trapezoidal rule in t,
    // which converges geometrically for this integrand).
    static double k1e(double z) {
        const int n = 4000;
        const double tmax = std::acosh(1.0 + 60.0 / z) + 1.0;
        const double h = tmax / n;
        double s = 0.5;  // t = 0 term: exp(0) cosh(0) / 2
        for (int i = 1; i <= n; i++) {
            double t = i * h;
            s += std::exp(-z * (std::cosh(t) - 1.0)) * std::cosh(t);
        }
        return s * h;
    }

    void finalize() {
        if (kind == SCN_LASER) {
            // plateau density given as a fraction of the critical density of
            // the laser wavelength: n_c = eps0 m omega0^2 / e^2
            double w0 = 2.0 * M_PI * vx_cs / lambda0;
            n_ref = n_over_nc * vx_eps0 * mass * w0 * w0 / (e_charge * e_charge);
        }
        omega_p = std::sqrt(n_ref * e_charge * e_charge / (vx_eps0 * mass));
        unit_len = vx_cs / omega_p;
        unit_p = mass * vx_cs;
        norm = 1.0 / (2.0 * k1e(1.0 / theta));
        // portable seeded phases: raw 64-bit mt19937 outputs -> [0,1)
        std::mt19937_64 gen(seed);
        for (int m = 0; m < 64; m++) {
            uint64_t u = gen();
            phases[m] = 2.0 * M_PI * (double)(u >> 11) * 0x1.0p-53;
        }
    }

    double dx() const { return length * unit_len / nx; }
    double dp() const { return (pmax - pmin) * unit_p / np; }
    double pmin_si() const { return pmin * unit_p; }

    // 1D Maxwell-Juttner in normalised momentum u, shifted to drift ud,
    // normalised to unit density: exp(-(gamma-1)/theta) / (2 K1e(1/theta)).
    double mj(double u, double ud) const {
        double v = u - ud;
        double g = std::sqrt(1.0 + v * v);
        return norm * std::exp(-(g - 1.0) / theta);
    }

    // f(x, p) in SI units (m^-3 per kg m s^-1): Settings::InitialDistribution
    // hook (veritas.cpp:107-115 shows the reference's own example).
    double initial(double x, double p, int species) const {
        (void)species;
        double xn = x / unit_len, u = p / unit_p;
        double n = 0.0, fu = 0.0;
        if (kind == SCN_LANDAU) {
            n = 1.0 + amplitude * std::cos(k * xn);
            fu = mj(u, 0.0);
        } else if (kind == SCN_TWO_STREAM) {
            double pert = 0.0;
            for (int m = 1; m <= modes; m++)
                pert += std::cos(m * 2.0 * M_PI * xn / length + phases[m - 1]);
            n = 1.0 + amplitude * pert;
            fu = 0.5 * mj(u, drift) + 0.5 * mj(u, -drift);
        } else {  // laser: underdense plasma with a linear ramp then a plateau
            double lam = lambda0 / unit_len;  // wavelength in c/omega_p
            double start = 10.0 * lam;        // vacuum gap before the ramp
            double xr = (xn - start) / (ramp * lam);
            double prof = xr <= 0 ? 0.0 : (xr < 1 ? xr : 1.0);
            if (gap > 0.0) {  // two slabs: vacuum of `gap` wavelengths mid-plateau
                double mid = 0.5 * (start + ramp * lam + length);
                if (std::fabs(xn - mid) < 0.5 * gap * lam) prof = 0.0;
            }
            n = prof;
            fu = mj(u, 0.0);
        }
        return n_ref * n * fu / unit_p;
    }

    // Transverse magnetic drive at x = 0 (GetBY/GetBZ hooks,
    // veritas.cpp:80-104): zero except for the laser scenario, where a
    // linearly polarised (y) Gaussian pulse with peak a0 enters from the left.
    double laser_env(double t) const {
        double T = lambda0 / vx_cs;
        double fw = fwhm_cycles * T;
        double tc = 2.0 * fw;
        double s = fw / (2.0 * std::sqrt(2.0 * std::log(2.0)));
        return std::exp(-(t - tc) * (t - tc) / (2 * s * s));
    }
    double by(double x, double t) const { (void)x; (void)t; return 0.0; }
    double bz(double x, double t) const {
        if (kind != SCN_LASER) return 0.0;
        double k0 = 2.0 * M_PI / lambda0, w0 = vx_cs * k0;
        double Ec = mass * vx_cs * w0 / e_charge;  // field for a0 = 1
        double E = a0 * Ec * laser_env(t - x / vx_cs) * std::sin(w0 * t - k0 * x);
        return E / vx_cs;
    }

    // The scenario as the reference's inputs: Input and Particles
    // (Settings.hpp:4-31) with one species, and the user hooks pointing at
    // this scenario (which must outlive their use).
    vx_settings_desc desc() const {
        vx_settings_desc d{};
        d.minEfficiency = min_efficiency;
        d.dx = dx();
        d.refinementCriteria = refinement_criteria;
        d.cfl = cfl;
        d.sizeWeight = size_weight;
        d.preLength = pre_length;
        d.postLength = post_length;
        d.nx = (uint32_t)nx;
        d.r = (uint32_t)r;
        d.Lfinest = (uint32_t)levels;
        d.quadratureDepth = quadrature_depth;
        d.n_species = 1;
        d.mass[0] = mass;
        d.charge[0] = charge;
        d.dp[0] = this->dp();
        d.pmin[0] = pmin_si();
        d.np[0] = (uint32_t)np;
        return d;
    }
    static double hook_initial(void *u, double x, double p, int32_t sp) {
        return ((const Scenario *)u)->initial(x, p, sp);
    }
    static double hook_by(void *u, double x, double t) { return ((const Scenario *)u)->by(x, t); }
    static double hook_bz(void *u, double x, double t) { return ((const Scenario *)u)->bz(x, t); }
    vx_hooks hooks() const {
        vx_hooks h{};
        h.user = (void *)this;
        h.initial_distribution = &hook_initial;
        h.get_by = &hook_by;
        h.get_bz = &hook_bz;
        h.refinement_override = nullptr;  // the built-in scenarios define none
        h.settings_override = nullptr;
        return h;
    }

    static Scenario from_params(const vx_scenario_params &c) {
        Scenario s;
        s.kind = c.kind;
        s.nx = c.nx;
        s.np = c.np;
        s.levels = c.levels;
        s.r = c.r;
        s.quadrature_depth = c.quadrature_depth;
        s.modes = c.modes;
        s.seed = c.seed;
        s.cfl = c.cfl;
        s.min_efficiency = c.min_efficiency;
        s.refinement_criteria = c.refinement_criteria;
        s.size_weight = c.size_weight;
        s.n_ref = c.n_ref;
        s.length = c.length;
        s.pmin = c.pmin;
        s.pmax = c.pmax;
        s.theta = c.theta;
        s.amplitude = c.amplitude;
        s.k = c.k;
        s.drift = c.drift;
        s.a0 = c.a0;
        s.n_over_nc = c.n_over_nc;
        s.ramp = c.ramp;
        s.fwhm_cycles = c.fwhm_cycles;
        s.lambda0 = c.lambda0;
        s.gap = c.gap;
        s.finalize();
        s.pre_length = c.pre_length * s.unit_len;
        s.post_length = c.post_length * s.unit_len;
        return s;
    }
};

}  // namespace vx
