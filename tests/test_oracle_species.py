"""More than one species and user hooks, on the CPU.

The reference advances one Mesh per species (SolverManager.cpp:12-16, 33-35)
and sums their charge and current (EMSolver.cpp:104-122); its inputs and its
physics come from user code (Input/Particles and the hooks of
Settings.hpp:45-64, veritas.cpp:7-115).  Here:

- hooks written in Python reproduce a built-in scenario bit for bit;
- the assembled charge is the sum of the species' charges, each species'
  charge is the one it has alone, and a species of the opposite charge
  contributes the opposite charge;
- a cold plasma oscillates at sqrt(sum_s omega_ps^2): sqrt(2) omega_pe for a
  pair plasma, sqrt(1 + 1/4) omega_pe with positive ions of mass 4 m_e, which
  pins each species' mass and charge in the moments and the force;
- the host library's Settings (no GPU) applies settingsOverride before its
  checks, rejects bad inputs with the reference's messages and prints the
  reference's banner (Settings.cpp:5-125), which the reference's viewer parses
  (viewers/compositeViewer.m:10-24)."""
import ctypes
import math
import re

import numpy as np
import pytest

import veritas_amd as va
from veritas_amd import params as P
from _oracle import OracleRun, lib

N0 = 1.0e24
W_P = math.sqrt(N0 * P.Q_E * P.Q_E / (P.EPS0 * P.M_E))
UNIT_X = P.C / W_P
MC = P.M_E * P.C


def landau_hooks(p):
    """The built-in Landau scenario's inputs and InitialDistribution
    (veritas_amd/csrc/scenario.h), restated in Python with the same
    operations in the same order."""
    m, e, n_ref = P.M_E, P.Q_E, p.n_ref
    omega_p = math.sqrt(n_ref * e * e / (P.EPS0 * m))
    unit_len = P.C / omega_p
    unit_p = m * P.C
    norm = 1.0 / (2.0 * lib().vxo_k1e(1.0 / p.theta))
    d = P.SettingsDesc.make(
        nx=p.nx, dx=p.length * unit_len / p.nx, cfl=p.cfl, Lfinest=p.levels, r=p.r,
        minEfficiency=p.min_efficiency, refinementCriteria=p.refinement_criteria,
        sizeWeight=p.size_weight, quadratureDepth=p.quadrature_depth,
        species=[dict(mass=m, charge=-e, np=p.np, dp=(p.pmax - p.pmin) * unit_p / p.np,
                      pmin=p.pmin * unit_p)])

    def f0(x, pp, sp):
        xn, u = x / unit_len, pp / unit_p
        n = 1.0 + p.amplitude * math.cos(p.k * xn)
        g = math.sqrt(1.0 + u * u)
        return n_ref * n * (norm * math.exp(-(g - 1.0) / p.theta)) / unit_p

    return d, P.Hooks(f0)


def test_python_hooks_reproduce_the_builtin_scenario():
    p = P.landau(nx=32, np=64)
    d, h = landau_hooks(p)
    a, b = OracleRun(p), OracleRun.from_desc(d, h)
    assert np.array_equal(a.get("f1"), b.get("f1"))
    dt = a.calc_dt()
    assert dt == b.calc_dt()
    a.advance(dt, 4)
    b.advance(dt, 4)
    assert np.array_equal(a.get("f1"), b.get("f1"))
    assert np.array_equal(a.get("E"), b.get("E"))


def maxwellian(theta, drift=lambda x: 0.0, species=(0,), n0=N0):
    def f0(x, p, sp):
        if sp not in species:
            return n0 * math.exp(-0.5 * (p / MC) ** 2 / theta) / math.sqrt(2 * math.pi * theta) / MC
        u = p / MC - drift(x)
        return n0 * math.exp(-0.5 * u * u / theta) / math.sqrt(2 * math.pi * theta) / MC
    return f0


def cold_desc(masses, charges, nx=64, n_p=64, theta=1e-4, eps=0.01, k=0.5):
    L = 2 * math.pi / k * UNIT_X
    pm = (8 * math.sqrt(theta) + eps) * MC
    sp = [dict(mass=m * P.M_E, charge=q * P.Q_E, np=n_p, dp=2 * pm / n_p, pmin=-pm)
          for m, q in zip(masses, charges)]
    return P.SettingsDesc.make(nx=nx, dx=L / nx, species=sp, cfl=0.4), 2 * math.pi / L


def test_species_charges_add_up():
    """Electrons, positrons and an electron-only run from the same data: the
    total is the ordered sum (EMSolver.cpp:104-122), the electrons' charge is
    the one they have alone, the positrons' its negative."""
    d2, k = cold_desc([1.0, 1.0], [-1.0, 1.0])
    d1, _ = cold_desc([1.0], [-1.0])
    h = P.Hooks(maxwellian(1e-4, lambda x: 0.01 * math.sin(k * x), species=(0, 1)))
    two, one = OracleRun.from_desc(d2, h), OracleRun.from_desc(d1, h)
    two.moments()
    one.moments()
    c0, c1 = two.get("charge_sp0"), two.get("charge_sp1")
    assert np.array_equal(two.get("charge"), (0.0 + c0) + c1)
    assert np.array_equal(c0, one.get("charge"))
    assert np.array_equal(c1, -c0)
    assert np.abs(c0).min() > 0


def oscillation_frequency(masses, charges, t_end=25.0):
    """A velocity perturbation eps sin(kx) of the electrons only (species 0)
    of a cold plasma (theta = 1e-4, k c/omega_pe = 0.5): E ~ sin(kx) oscillates
    at the plasma frequency of all species (Bohm-Gross corrections are below
    1e-4 here).  Returns omega / omega_pe from the zero crossings."""
    d, k = cold_desc(masses, charges)
    o = OracleRun.from_desc(d, P.Hooks(maxwellian(1e-4, lambda x: 0.01 * math.sin(k * x))))
    dt = o.calc_dt(0.4)
    xs = (np.arange(d.nx) + 0.5) * d.dx
    ts, sig = [], []
    for _ in range(int(t_end / W_P / dt)):
        o.advance(dt, 1)
        sig.append(2 * np.mean(o.get("E") * np.sin(k * xs)))
        ts.append(o.scalars()["time"] * W_P)
    z = [ts[i] - sig[i] * (ts[i + 1] - ts[i]) / (sig[i + 1] - sig[i])
         for i in range(len(sig) - 1) if sig[i] * sig[i + 1] < 0]
    assert len(z) >= 6
    return math.pi / np.mean(np.diff(z))


def test_plasma_frequency_of_every_species_enters():
    w_e = oscillation_frequency([1.0], [-1.0])
    assert abs(w_e - 1.0) < 2e-3, w_e
    w_pair = oscillation_frequency([1.0, 1.0], [-1.0, 1.0])
    assert abs(w_pair / w_e - math.sqrt(2.0)) < 3e-3 * math.sqrt(2.0), (w_pair, w_e)
    w_ion = oscillation_frequency([1.0, 4.0], [-1.0, 1.0])
    assert abs(w_ion / w_e - math.sqrt(1.25)) < 3e-3 * math.sqrt(1.25), (w_ion, w_e)


# ---------------------------------------------------------------- host Settings (no GPU)

def _desc_two_species():
    d, _ = cold_desc([1.0, 1836.0], [-1.0, 1.0], nx=76, n_p=50)
    return d


def test_settings_override_runs_before_the_checks():
    """Settings.cpp:31: settingsOverride sees the copied inputs and may change
    them; the checks of Settings.cpp:35-50 apply to the result."""
    d = _desc_two_species()
    d.nx = 75  # not divisible by r = 2: rejected unless the override fixes it
    h = P.Hooks(maxwellian(1e-4), settings_override=lambda s: setattr(s, "nx", 76))
    text = va.settings_banner(d, h)
    assert re.search(r"nx:\s+76\b", text)
    with pytest.raises(va.VeritasError, match="'x_size' must be completely divisible"):
        va.settings_banner(d, P.Hooks(maxwellian(1e-4)))
    d = _desc_two_species()
    d.np[1] = 51
    with pytest.raises(va.VeritasError, match=r"'p_size' at index \(1\) must be completely divisible"):
        va.settings_banner(d, P.Hooks(maxwellian(1e-4)))
    d = _desc_two_species()
    d.r = 3
    with pytest.raises(va.VeritasError, match="must be an even value"):
        va.settings_banner(d, P.Hooks(maxwellian(1e-4)))


def test_banner_carries_what_the_viewer_parses():
    """viewers/compositeViewer.m:10-24 reads nx, dx, the refinement ratio and
    each species' np and dp from the banner (Settings.cpp:60-125)."""
    d = _desc_two_species()
    text = va.settings_banner(d, P.Hooks(maxwellian(1e-4)))
    lines = text.splitlines()
    assert all(len(l) <= 80 for l in lines)
    assert " Configuration Parameters " in text and " Particle Properties " in text
    assert float(re.search(r"nx:\s+([0-9.]+)", text).group(1)) == 76
    assert float(re.search(r"dx:\s+([0-9.eE+-]+)", text).group(1)) == pytest.approx(d.dx, rel=1e-4)
    assert float(re.search(r"ratio:\s+([0-9.]+)", text).group(1)) == 2
    assert [int(v) for v in re.findall(r"np:\s+([0-9]+)", text)] == [50, 50]
    dps = [float(v) for v in re.findall(r"dp:\s+([0-9eE+.-]+)", text)]
    assert dps == pytest.approx([d.dp[0], d.dp[1]], rel=1e-4)
    assert text.count("Mass:") == 2 and text.count("-" * 80) == 1
