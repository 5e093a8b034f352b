"""GPU parity with more than one species and with user hooks.

The reference advances one Mesh per species and sums their charge and current
(SolverManager.cpp:12-16, 33-35; EMSolver.cpp:104-122); its inputs and hooks
are user code (Settings.hpp:45-64; veritas.cpp:7-115; its own example runs
electrons and ions, veritas.cpp:15-19).  Each case runs the GPU library and
the CPU oracle from the same SettingsDesc and Python hooks and compares them
bit for bit."""
import math
import sys

import numpy as np
import pytest

import veritas_amd as va
from veritas_amd import params as P
from _oracle import OracleRun
from test_gpu_parity import assert_same, assert_ulp
from test_oracle_species import MC, UNIT_X, cold_desc, landau_hooks, maxwellian

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare_species(g, o, n_species, tag, what=("f1", "f0")):
    for sp in range(n_species):
        lay_g, lay_o = g.layout(sp), o.layout(sp)
        assert np.array_equal(lay_g, lay_o), f"{tag}: species {sp} layouts differ\n{lay_g}\n{lay_o}"
        for lvl in sorted(set(lay_o[:, 0])):
            for k in range(int((lay_o[:, 0] == lvl).sum())):
                for w in what:
                    assert_same(f"{tag} species {sp} level {lvl} patch {k} {w}",
                                g.patch(lvl, k, w, ghosts=2, sp=sp), o.patch(lvl, k, w, ghosts=2, sp=sp))


def test_python_hooks_reproduce_the_builtin_scenario():
    p = P.landau(nx=32, np=64)
    d, h = landau_hooks(p)
    say("case: Landau 32x64 from Python hooks vs the built-in scenario, 4 steps")
    a, b, o = va.Run(p), va.Run.from_desc(d, h), OracleRun(p)
    dt = o.calc_dt()
    for r in (a, b):
        assert r.calc_dt() == dt
        r.advance(dt, 4)
    o.advance(dt, 4)
    assert_same("scenario f1", a.get("f1"), o.get("f1"))
    assert_same("hooks f1", b.get("f1"), o.get("f1"))


@pytest.mark.parametrize("fused", [True, False])
def test_electrons_and_ions_match_the_oracle(fused):
    """Electrons with a velocity perturbation and ions of mass 1836 m_e
    (veritas.cpp:15-19), 20 steps; both species' f, their charges and E."""
    d, k = cold_desc([1.0, 1836.0], [-1.0, 1.0], nx=64, n_p=64)
    h = P.Hooks(maxwellian(1e-4, lambda x: 0.01 * math.sin(k * x)))
    say(f"case: electrons + ions 64x64, fused={fused}, 20 steps")
    g, o = va.Run.from_desc(d, h, fused=fused), OracleRun.from_desc(d, h)
    dt = o.calc_dt(0.4)
    assert g.calc_dt(0.4) == dt
    g.advance(dt, 20)
    o.advance(dt, 20)
    compare_species(g, o, 2, "e+i")
    g.stage_moments()
    o.moments()
    for w in ("charge_sp0", "charge_sp1", "charge", "J"):
        assert_same(w, g.get(w), o.get(w))
    assert_ulp("E", g.get("E"), o.get("E"))
    assert np.abs(o.get("E")).max() > 0


def test_two_species_hierarchies_with_a_refinement_override():
    """Two species on three levels each; the RefinementOverride hook
    (Settings.hpp:63, veritas.cpp:76-78) flags a box of phase space, the
    error estimate alone never would (criterion 10).  Layouts, every patch's f
    after steps and after a regrid, and the progress table's rows."""
    d, k = cold_desc([1.0, 4.0], [-1.0, 1.0], nx=32, n_p=16, theta=4e-3, eps=0.05)
    # quadrature of the initial data with r^depth points per cell and axis (a
    # Python hook is called for each)
    d.Lfinest, d.refinementCriteria, d.minEfficiency, d.quadratureDepth = 3, 10.0, 0.7, 0
    L = d.nx * d.dx

    def override(x, p, depth, sp):
        return 0.3 * L < x < (0.55 if sp == 0 else 0.45) * L and abs(p) < 0.12 * MC * (2 if sp else 1)

    h = P.Hooks(maxwellian(4e-3, lambda x: 0.05 * math.sin(k * x)), refinement_override=override)
    say("case: two species, 3 levels, refinement override, 6 steps + regrid + 4 steps")
    g, o = va.Run.from_desc(d, h), OracleRun.from_desc(d, h)
    for sp in range(2):
        assert len(o.layout(sp)) >= 3, o.layout(sp)
    compare_species(g, o, 2, "set-up", what=("f1",))
    dt = o.calc_dt(0.4)
    assert g.calc_dt(0.4) == dt
    g.advance(dt, 6)
    o.advance(dt, 6)
    compare_species(g, o, 2, "6 steps")
    g.regrid(6 * dt)
    o.regrid()
    compare_species(g, o, 2, "regrid")
    g.advance(dt, 4)
    o.advance(dt, 4)
    compare_species(g, o, 2, "4 more steps")
    # SolverManager.cpp:9-22, 48-57, 73-100: header, set-up row, one row per regrid,
    # a line per species with its used (allocated) levels and patches per level
    screen = g.screen().splitlines()
    assert screen[0].startswith("--- Time (T) ----|--- Particle ---|")
    assert len(screen) == 1 + 2 * 2
    for row, sp in zip(screen[1:], (0, 1, 0, 1)):
        assert f" {sp} " in row and "(3)" in row, row
