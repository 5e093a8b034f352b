"""GPU parity of the field-only phase (SolverManager::AdvanceFields,
SolverManager.cpp:41-46) followed by full steps, as the reference's driver runs
a laser case (veritas.cpp:129-144): dt = T/400 of field-only steps while
t < 3T, then Advance with dt = min(CalculateDt(cfl), T/400)."""
import sys

import numpy as np
import pytest

import veritas_amd as va
from _amr_cases import two_slab_laser
from _oracle import OracleRun
from test_gpu_parity import assert_same, assert_ulp

pytestmark = pytest.mark.gpu

C = 299792458.0


def test_vacuum_lead_in_then_steps_match_the_oracle():
    p = two_slab_laser(levels=1, fwhm_cycles=2.0)
    print("case: laser 128x64, 3T of AdvanceFields at T/400, then 10 adaptive steps", flush=True)
    sys.stdout.flush()
    g, o = va.Run(p), OracleRun(p)
    T = p.lambda0 / C
    dt = T / 400
    t = dt
    lead = 0
    while t <= 3 * T:  # veritas.cpp:139-144
        lead += 1
        t += dt
    g.advance_fields(dt, lead)
    o.advance_fields(dt, lead)
    for w in ("By", "Bz", "Ey", "Ez", "Ay", "Az", "a2"):
        assert_same(f"lead-in {w}", g.get(w), o.get(w))
    assert np.abs(o.get("Ey")).max() > 0 and np.abs(o.get("a2")).max() > 0
    assert_same("lead-in f1 untouched", g.get("f1"), o.get("f1"))
    assert g.scalars()["time"] == o.scalars()["time"]
    for _ in range(10):  # veritas.cpp:134-141
        step = min(o.calc_dt(p.cfl), dt)
        assert min(g.calc_dt(p.cfl), dt) == step
        g.advance(step)
        o.advance(step)
    assert_same("f1", g.get("f1"), o.get("f1"))
    for w in ("By", "Bz", "Ey", "Ez", "Ay", "Az", "a2"):
        assert_same(w, g.get(w), o.get(w))
    assert_ulp("E", g.get("E"), o.get("E"))
    g.close()
    o.close()
