"""Pins of the oracle's Vlasov advance (oracle/oracle.cpp, the restatement of
Rectangle::FCTTimeStep, Rectangle.cpp:1255-1623, driven as in Mesh::Advance,
Mesh.cpp:64-89) by exact solutions of the Vlasov equation.

The GPU path is bit-compared with the oracle, so a slip in the oracle's flux
formula, advection speeds or Runge-Kutta combination would be reproduced on
the GPU and pass every parity test.  These pins fix those parts by what the
mathematics gives, with the fields imposed (the potential zeroed and Ex0 set
before every stage, A = 0) so that the exact solution is known:

- free streaming, E = 0: f(x, p, t) = f0(x - v(p) t, p) with v = p / (m gamma);
  exact cell averages come from an erf in x and Gauss-Legendre in p;
- f linear in x: the flux differences are the same at every stage, so the
  result must not depend on dt, and the only error left is the p-average of
  v f in the x-flux, which the 1/48 product term makes fourth order in dp
  (Rectangle.cpp:1330-1345);
- a uniform force on x-uniform f: f(p, t) = f0(p - q E t);
- a force E0 cos(omega t) on f linear in p: the semi-discrete system is then a
  quadrature of E over the stages, whose error falls at the order of the
  tableau and its stage clock (b and c of Rectangle.cpp:1396-1517 and
  Settings.cpp:166-179 integrate polynomials to degree 3 exactly).

Each is chosen so that a plausible slip fails one of them: an index of the RK
flux combination (FxH[k] -> FxH[k-1]), the sign of the 1/48 term, a factor in
the advection speed `ex` or in the force `ep`, an RK coefficient
(tests/test_oracle_plants.py plants each in a copy of the oracle).

Patches have closed walls with zero ghosts (Rectangle.cpp:1257-1260,
BoundaryCondition.cpp:6-8), so every profile is kept off the walls or only
cells out of the walls' reach are compared."""
import math

import numpy as np
from numpy.polynomial.legendre import leggauss
from scipy.special import erf

from _oracle import OracleRun
from veritas_amd import params as P

C = 299792458.0          # veritas.hpp:25 (cs)
M_E = 9.10938291e-31     # veritas.cpp:15
Q_E = -1.60217657e-19    # veritas.cpp:16, electrons
EPS0 = 8.854187817e-12   # veritas.hpp:23
N_REF = 1.0e24
W_P = math.sqrt(N_REF * Q_E ** 2 / (EPS0 * M_E))  # plasma frequency of n_ref
UNIT_X = C / W_P                                   # c / omega_p in metres
MC = M_E * C

_G, _W = leggauss(8)


def p_nodes(n_p, dp, pmin):
    """Gauss-Legendre nodes (8 per cell) of every p cell, shape (n_p, 8)."""
    lower = pmin + np.arange(n_p) * dp
    return lower[:, None] + (_G[None, :] + 1.0) * 0.5 * dp


def p_average(values):
    return (values * _W).sum(-1) * 0.5


def velocity(p):
    u = p / MC
    return p / (M_E * np.sqrt(1.0 + u * u))


def run(nx, n_p, length=20.0, plen=0.9):
    """An oracle run on [0, length) c/omega_p x [-plen, plen] m_e c; the
    scenario's own initial data are replaced by the caller."""
    p = P.landau(nx=nx, np=n_p, length=length, pmin=-plen, pmax=plen, amplitude=0.0)
    o = OracleRun(p)
    s = o.scalars()
    return o, s["dx"], s["dp"], s["pmin"]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- free streaming

def pulse_averages(nx, n_p, dx, dp, pmin, t, x0, sig, ps):
    """Exact cell averages of f0(x - v t, p), f0 = exp(-(x-x0)^2/2sig^2) *
    exp(-u^2/2ps^2), u = p / (m c)."""
    pq = p_nodes(n_p, dp, pmin)
    shift = velocity(pq) * t
    prof = np.exp(-0.5 * (pq / MC / ps) ** 2)
    edges = np.arange(nx + 1) * dx

    def anti(x):  # antiderivative of the x profile
        return sig * math.sqrt(math.pi / 2) * erf((x - x0) / (sig * math.sqrt(2)))

    out = np.empty((nx, n_p))
    for i in range(nx):
        xi = (anti(edges[i + 1] - shift) - anti(edges[i] - shift)) / dx
        out[i] = p_average(xi * prof)
    return out


def free_streaming_error(nx, n_p, t_end=2.0, cfl=0.4):
    o, dx, dp, pmin = run(nx, n_p)
    L = nx * dx
    o.set_f(pulse_averages(nx, n_p, dx, dp, pmin, 0.0, 0.5 * L, L / 40, 0.12))
    T = t_end / W_P
    steps = int(math.ceil(T / (cfl * dx / C)))
    o.vlasov_imposed(T / steps, steps)
    t = o.scalars()["time"]
    return rel_l2(o.get("f1"), pulse_averages(nx, n_p, dx, dp, pmin, t, 0.5 * L, L / 40, 0.12))


def test_free_streaming_pulse_converges_at_third_order():
    """A Gaussian pulse (sigma = L/40, u-width 0.12) free-streams for
    2/omega_p; the p-dependent speed shears it across every p cell.  dx, dp
    and dt are halved together (dt = 0.4 dx / c): the error against the exact
    cell averages falls at third order (the two-candidate WENO faces), an
    inconsistent combination or speed would not converge."""
    errs = [free_streaming_error(nx, n_p) for nx, n_p in ((64, 32), (128, 64), (256, 128))]
    orders = [math.log2(a / b) for a, b in zip(errs, errs[1:])]
    assert errs[-1] < 4e-4, errs
    assert min(orders) > 2.7, (errs, orders)


def linear_x_run(n_p, cfl, nx=64, t_end=2.0, ps=0.12):
    """f0 = (1 + x/L) exp(-u^2/2ps^2), E = 0.  Exact: f = (1 + (x - v t)/L) P(p),
    whose cell averages are (1 + x_i/L) <P>_j - (t/L) <v P>_j."""
    o, dx, dp, pmin = run(nx, n_p)
    L = nx * dx
    pq = p_nodes(n_p, dp, pmin)
    prof = np.exp(-0.5 * (pq / MC / ps) ** 2)
    xc = (np.arange(nx) + 0.5) * dx
    avg_p = p_average(prof)[None, :]
    avg_vp = p_average(velocity(pq) * prof)[None, :]

    def exact(t):
        return (1.0 + xc[:, None] / L) * avg_p - (t / L) * avg_vp

    o.set_f(exact(0.0))
    T = t_end / W_P
    steps = int(math.ceil(T / (cfl * dx / C)))
    o.vlasov_imposed(T / steps, steps)
    # the walls' disturbance travels at most c t = 2 c/omega_p (7 cells) in
    # this time: compare the middle 16 columns
    mid = slice(nx // 2 - 8, nx // 2 + 8)
    return o.get("f1")[mid], exact(o.scalars()["time"])[mid]


def test_linear_profile_is_exact_in_time_and_fourth_order_in_dp():
    """On f linear in x the high-order x-flux differences between neighbouring
    faces are the same at every stage (the profile drifts, its slope stays),
    so the step is exact in time: dt must not matter.  What is left is the
    p-average of v f through the x-face, approximated by
    f_face * <v> + (1/48) (f_{j+1} - f_{j-1}) (v_{j+1} - v_{j-1}), which is
    fourth order in dp (a flipped 1/48 term leaves second order, a wrong
    speed factor no convergence)."""
    a, ex = linear_x_run(64, 0.4)
    b, _ = linear_x_run(64, 0.2)
    assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()
    errs = []
    for n_p in (32, 64, 128):
        f, fe = linear_x_run(n_p, 0.4)
        errs.append(rel_l2(f, fe))
    ratios = [a / b for a, b in zip(errs, errs[1:])]
    assert errs[-1] < 1e-7, errs
    assert min(ratios) > 12.0, (errs, ratios)


# ---------------------------------------------------------------- uniform force

def force_run(n_p, E0, omega, steps, init, nx=32, length=20.0, plen=1.2):
    o, dx, dp, pmin = run(nx, n_p, length=length, plen=plen)
    o.set_f(np.broadcast_to(init(pmin, dp)[None, :], (nx, n_p)))
    return o, dx, dp, pmin


def gaussian_force_error(n_p, t_end=2.0, u_shift=0.2, ps=0.12, nx=32):
    """x-uniform f = exp(-u^2/2ps^2) under a uniform E that moves it by
    u_shift in t_end; dt by the reference's CFL rule (c/dx + |qE|/dp)."""
    T = t_end / W_P
    E0 = u_shift * MC / (Q_E * T)

    def exact(pmin, dp, t):
        S = Q_E * E0 * t / MC
        return p_average(np.exp(-0.5 * ((p_nodes(n_p, dp, pmin) / MC - S) / ps) ** 2))

    o, dx, dp, pmin = force_run(n_p, E0, 0.0, 0, lambda pm, d: exact(pm, d, 0.0), nx=nx)
    steps = int(math.ceil(T / (0.4 / (C / dx + abs(Q_E * E0) / dp))))
    o.vlasov_imposed(T / steps, steps, E0)
    mid = slice(nx // 2 - 4, nx // 2 + 4)  # x walls: out of reach (2.4 cells)
    f = o.get("f1")[mid]
    return rel_l2(f, np.broadcast_to(exact(pmin, dp, o.scalars()["time"]), f.shape))


def test_uniform_force_shifts_p_at_the_exact_rate():
    """dp/dt = q E: a Gaussian in p moves by q E t (an E-force sign or factor
    slip moves it elsewhere); the error falls with dp at better than second
    order (third-order faces, clipped at the peak by the limiter)."""
    errs = [gaussian_force_error(n_p) for n_p in (64, 128, 256)]
    orders = [math.log2(a / b) for a, b in zip(errs, errs[1:])]
    assert errs[-1] < 1e-3, errs
    assert min(orders) > 2.0, (errs, orders)


def oscillating_force_error(steps, n_p=128, nx=64, length=40.0, plen=1.2, frac=0.37):
    """f = 1 + (u - S(t))/2 with S(t) = q E0 sin(omega t) / (omega m c), one dp of
    amplitude, period 2/omega_p, stopped at 0.37 of the period; the error is
    taken over the middle of the domain, out of reach of the walls."""
    period = 2.0 / W_P
    omega = 2 * math.pi / period
    o, dx, dp, pmin = run(nx, n_p, length=length, plen=plen)
    E0 = dp * omega / Q_E
    uc = (pmin + (np.arange(n_p) + 0.5) * dp) / MC

    def exact(t):
        S = Q_E * E0 * math.sin(omega * t) / omega / MC
        return np.broadcast_to((1.0 + 0.5 * (uc - S))[None, :], (nx, n_p))

    o.set_f(exact(0.0))
    T = frac * period
    o.vlasov_imposed(T / steps, steps, E0, omega)
    d = (o.get("f1") - exact(o.scalars()["time"]))[nx // 2 - 8:nx // 2 + 8, n_p // 4:3 * n_p // 4]
    return float(np.abs(d).max())


def test_time_dependent_force_converges_at_the_order_of_the_tableau():
    """With E0 cos(omega t) evaluated at the stage clock t_n + c_i dt, the
    update of a linear profile is the quadrature sum_i b_i E(t_n + c_i dt) dt of
    the impulse: fourth order, error ratio 16 per halving of dt.  (With fewer
    than 32 steps per 0.37 period the limiter acts, because the predictor uses
    only the stage-0 low-order flux, Rectangle.cpp:1405-1511, and the order
    drops to one; that is the reference's method, not tested here.)"""
    e32, e64 = oscillating_force_error(32), oscillating_force_error(64)
    assert e32 > 1e-12  # above rounding, so the ratio means something
    assert 12.0 < e32 / e64 < 20.0, (e32, e64)
