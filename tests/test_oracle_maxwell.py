"""Pins of the oracle's transverse Maxwell solver (oracle/oracle.cpp, the
restatement of EMSolver.cpp:194-338 and 479-619) by what the mathematics fixes.

The reference ships no vectors for its field solver, so the pins are:
- a vacuum pulse driven at the left edge through the GetBZ hook
  (veritas.cpp:80-104; EMSolver.cpp:492-499) must equal the travelling wave
  E_y(x, t) = c B_z(0, t - x/c) (B_z and A_y = -int E_y dt alike), with E_y at
  x = (i + 1/2) dx and B_z at x = (i + 1) dx of the padded index i;
- the error falls with the grid faster than second order (4th-order bulk
  stencil, 2nd-order pads) and with the step at exactly fourth order (the
  six-stage tableau with the stage clock of Settings.cpp:166-179);
- a weak pulse in a uniform cold plasma travels at the group velocity
  c sqrt(1 - n/n_c), which pins the current term -eps0_inv J A of
  EMSolver.cpp:525-531 (a dropped term gives c, a sign error a velocity above
  c, a factor of two 0.71 c at n = n_c/4).

Each catches a plausible mistake: a sign, an index shift of the staggering,
a stencil weight, an RK coefficient or stage time, the mu/eps factors."""
import math

import numpy as np

from _oracle import OracleRun
from veritas_amd import params as P

C = 299792458.0
M_E = 9.10938291e-31   # veritas.cpp:15-16
E_Q = 1.60217657e-19


def travelling_wave(p, t, x):
    """E_y of the right-going laser pulse that the drive injects (a0, FWHM of
    the field envelope in cycles, centred at twice the FWHM)."""
    lam = p.lambda0
    T = lam / C
    k0 = 2 * math.pi / lam
    w0 = C * k0
    Ec = M_E * C * w0 / E_Q
    fw = p.fwhm_cycles * T
    tc = 2 * fw
    s = fw / (2 * math.sqrt(2 * math.log(2)))
    tau = t - x / C
    return p.a0 * Ec * np.exp(-(tau - tc) ** 2 / (2 * s * s)) * np.sin(w0 * t - k0 * x)


def vacuum_run(nx, steps, t_end_cycles=6.0, what=("Ey",)):
    """Field-only steps (the plasma's J is zero after neutralisation) of an
    8-wavelength domain with a 2-cycle pulse, ending at t_end."""
    p = P.laser(nx=nx, np=8, levels=1, wavelengths=8.0, fwhm_cycles=2.0, pre_length=0.0,
                post_length=0.0)
    o = OracleRun(p)
    dt = t_end_cycles * (p.lambda0 / C) / steps
    o.advance_fields(dt, steps)
    s = o.get("scalars")
    out = {w: o.get(w) for w in what}
    o.close()
    ne = len(out[what[0]]) // 8  # 8 RK slots, slot 0 first (EMSolver.hpp:51-53)
    return p, s[1], s[2], int(s[7]), {w: v[:ne] for w, v in out.items()}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_vacuum_pulse_is_the_travelling_wave():
    p, t, dx, npre, f = vacuum_run(256, 3840, what=("Ey", "Bz", "Ay"))
    assert npre == 2
    ne = len(f["Ey"])
    i = np.arange(ne)
    inner = slice(0, ne - 4)  # the last two cells are frozen (EMSolver.cpp:539)
    ey = travelling_wave(p, t, (i + 0.5) * dx)
    assert rel_l2(f["Ey"][inner], ey[inner]) < 2e-3
    bz = travelling_wave(p, t, (i + 1.0) * dx) / C
    assert rel_l2(f["Bz"][inner], bz[inner]) < 2e-3
    tt = np.linspace(0.0, t, 20001)
    ay = -np.trapezoid(travelling_wave(p, tt[:, None], ((i + 0.5) * dx)[None, :]), tt, axis=0)
    assert rel_l2(f["Ay"][inner], ay[inner]) < 2e-3


def test_vacuum_pulse_converges_in_space():
    errs = []
    for nx in (128, 256, 512):
        steps = int(round(6 * (nx / 8) / 0.05))  # c dt / dx = 0.05: time error negligible
        p, t, dx, _, f = vacuum_run(nx, steps)
        ne = len(f["Ey"])
        ey = travelling_wave(p, t, (np.arange(ne) + 0.5) * dx)
        errs.append(rel_l2(f["Ey"][: ne - 4], ey[: ne - 4]))
    assert errs[0] < 2e-2
    assert errs[0] / errs[1] > 5.0 and errs[1] / errs[2] > 5.0, errs


def test_runge_kutta_is_fourth_order_in_time():
    ref = vacuum_run(256, int(round(6 * 32 / 0.025)))[4]["Ey"]
    errs = [rel_l2(vacuum_run(256, int(round(6 * 32 / c)))[4]["Ey"], ref) for c in (1.6, 0.8, 0.4)]
    for a, b in zip(errs, errs[1:]):
        assert 14.0 < a / b < 18.5, errs


def centroid_velocity(n_over_nc):
    """Velocity (units of c) of the |E_y|^2 centroid of a weak 5-cycle pulse
    inside a uniform plasma that starts 10 wavelengths from the drive."""
    waves, cpl = 44.0, 10
    p = P.laser(nx=int(waves * cpl), np=24, levels=1, wavelengths=waves, fwhm_cycles=5.0,
                a0=0.02, n_over_nc=n_over_nc, ramp=0.01, pmin=-0.15, pmax=0.15, theta=1e-3)
    o = OracleRun(p)
    T = p.lambda0 / C
    dt = o.calc_dt(0.4)
    steps = int(round(46 * T / dt))
    chunk = steps // 30
    track = []
    for _ in range(30):
        o.advance(dt, chunk)
        s = o.get("scalars")
        ey = o.get("Ey")
        ne = len(ey) // 8
        x = (np.arange(ne) + 0.5) * s[2] / p.lambda0
        w = np.where(x > 11.5, ey[:ne] ** 2, 0.0)
        if w.sum() > 0:
            track.append((s[1] / T, float(x @ w / w.sum())))
    o.close()
    track = np.array(track)
    m = (track[:, 1] > 20) & (track[:, 1] < 32)
    assert m.sum() >= 5
    return float(np.polyfit(track[m, 0], track[m, 1], 1)[0])


def test_plasma_current_term_gives_the_group_velocity():
    v_vac = centroid_velocity(1e-6)
    v_pl = centroid_velocity(0.25)
    assert abs(v_vac - 1.0) < 0.01
    # the same measurement in vacuum calibrates the centroid method
    assert abs(v_pl / v_vac - math.sqrt(1 - 0.25)) < 0.015 * math.sqrt(1 - 0.25)
