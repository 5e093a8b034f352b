"""Pins for the CPU oracle (oracle/oracle.cpp).

The reference ships no tests, fixtures or golden vectors, and cannot be built in
this image (it needs an mkl.h and a generated gitsha1.hpp).  Each oracle
function is therefore pinned by what the mathematics fixes: closed forms,
invariants, library routines (LAPACK via numpy, Bessel functions via scipy) and
brute force, chosen so that a dropped term, wrong sign or transposed operand
fails at least one check.  None of these re-types the oracle's formula.
"""
import math

import numpy as np
import pytest
import scipy.special as sp

from _oracle import OracleRun, lib
from veritas_amd.params import landau, two_stream

EPS = np.finfo(np.float64).eps


# ---------------------------------------------------------------- vx_log
def test_vx_log_within_one_ulp_of_libm():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-300, 300, 20000)),
                         1.0 + rng.uniform(-1e-3, 1e-3, 20000),
                         1.0 + rng.uniform(-1e-12, 1e-12, 5000),
                         [1.0, 2.0, 0.5, math.sqrt(2), 1e-300, 1e300]])
    L = lib()
    got = np.array([L.vxo_log(float(x)) for x in xs])
    ref = np.log(xs)
    ulps = np.abs(got.view(np.int64) - ref.view(np.int64))
    ulps[np.sign(got) != np.sign(ref)] = 0 if False else ulps[np.sign(got) != np.sign(ref)]
    assert np.all(ulps <= 1), (xs[ulps > 1], ulps.max())
    assert L.vxo_log(1.0) == 0.0


def test_scaled_bessel_k1e_matches_scipy():
    L = lib()
    for z in (1.0, 10.0, 50.0, 100.0, 200.0, 1000.0):
        assert L.vxo_k1e(z) == pytest.approx(sp.k1e(z), rel=1e-13)


# ---------------------------------------------------------------- WENO
def weno(f, right):
    return lib().vxo_weno_nomax(*map(float, f), int(right))


def test_weno_exact_for_linear_data():
    # cell averages of a linear function: both candidates are the exact face value
    for a, b in [(0.3, 1.7), (-2.0, 0.25), (1e-3, -5e-4)]:
        f = [a + b * k for k in range(1, 5)]
        for right in (True, False):
            assert weno(f, right) == pytest.approx(a + 2.5 * b, rel=1e-13, abs=1e-15)


def test_weno_constant_is_constant():
    for c in (0.0, 1.0, 3.5e-7, -2.0):
        for right in (True, False):
            assert weno([c] * 4, right) == pytest.approx(c, rel=1e-15, abs=0)


def test_weno_picks_the_upwind_candidate_at_a_jump():
    # left stencil (0,0,0) is smooth, right stencil (0,0,1) holds the jump
    f = [0.0, 0.0, 0.0, 1.0]
    # flow to +x: the larger weight goes to the left candidate -> ~0
    assert abs(weno(f, True)) < 1e-12
    # flow to -x: the larger weight goes to the right candidate -> fR = -1/6
    assert weno(f, False) == pytest.approx(-1.0 / 6.0, rel=1e-9)


def test_weno_mirror_symmetry():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        f = rng.uniform(0, 1, 4) * 10 ** rng.uniform(-3, 3)
        for right in (True, False):
            a = weno(f, right)
            b = weno(f[::-1], not right)
            assert a == pytest.approx(b, rel=1e-12, abs=1e-300)


def test_weno_smooth_data_third_order():
    # face value of cell averages of sin: error drops ~h^3 under refinement
    errs = []
    for h in (0.1, 0.05, 0.025):
        x = np.arange(1, 5) * h
        avg = (np.cos(x - h / 2) - np.cos(x + h / 2)) / h + 2.0
        exact = math.sin(2.5 * h) + 2.0
        errs.append(abs(weno(avg, True) - exact))
    assert errs[0] / errs[1] > 6 and errs[1] / errs[2] > 6


def test_face_weno_of_linear_and_symmetric():
    L = lib()
    assert L.vxo_face_weno(1.0, 2.0, 3.0, 4.0) == pytest.approx(2.5, rel=1e-14)
    assert L.vxo_face_weno(4.0, 3.0, 2.0, 1.0) == pytest.approx(2.5, rel=1e-14)
    assert L.vxo_face_weno(0.0, 0.0, 0.0, 0.0) == 0.0


# ---------------------------------------------------------------- Poisson
def dense_matrix(N):
    """The Poisson matrix exactly as EMSolver.cpp:41-67 assembles pM, read in
    column-major order as LAPACK does (A[r, c] = pM[c*N + r])."""
    w1, w2, w3 = -16.0 / 30.0 / 0.4, 1.0 / 30.0 / 0.4, 1 / 0.4
    pM = np.zeros(N * N)
    for i in range(2, N - 2):
        pM[i * N + i] = w3
        pM[i * N + i + 1] = w1
        pM[i * N + i - 1] = w1
        pM[i * N + i + 2] = w2
        pM[i * N + i - 2] = w2
    pM[0] = 1.0
    pM[N + 1] = w3
    pM[N + 2] = w1
    pM[N + 0] = w1
    pM[N + 3] = w2
    pM[2 * N - 1] = w2
    pM[(N - 2) * N + N - 2] = w3
    pM[(N - 2) * N + N - 1] = w1
    pM[(N - 2) * N + N - 3] = w1
    pM[(N - 2) * N + 0] = w2
    pM[(N - 2) * N + N - 4] = w2
    pM[(N - 1) * N + (N - 1)] = w3
    pM[(N - 1) * N + 0] = w1
    pM[(N - 1) * N + (N - 2)] = w1
    pM[(N - 1) * N + 1] = w2
    pM[(N - 1) * N + (N - 3)] = w2
    return pM.reshape(N, N).T.copy()


@pytest.mark.parametrize("nx", [64, 256, 1024])
def test_poisson_matches_lapack_gesv(nx):
    o = OracleRun(landau(nx=nx))
    A = dense_matrix(nx)
    rng = np.random.default_rng(nx)
    x = np.linspace(0, 2 * np.pi, nx, endpoint=False)
    cond = np.linalg.cond(A)
    Al = A.astype(np.longdouble)
    for b in (np.sin(3 * x) + 0.1 * np.cos(7 * x), rng.standard_normal(nx)):
        ref = np.linalg.solve(A, b)  # LAPACK dgesv = dgetrf + dgetrs, as EMSolver.cpp:75,177
        got = o.poisson(b)
        # agrees with plain LAPACK to LAPACK's own accuracy (cond * eps)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 10 * cond * EPS
        # and with LAPACK refined by an extended-precision residual to ~eps
        xr = ref.copy()
        for _ in range(3):
            r = (b.astype(np.longdouble) - Al @ xr.astype(np.longdouble)).astype(np.float64)
            xr = xr + np.linalg.solve(A, r)
        assert np.linalg.norm(got - xr) / np.linalg.norm(xr) < 20 * EPS


def test_poisson_refinement_is_idempotent_and_exact_on_solution():
    nx = 512
    o = OracleRun(landau(nx=nx))
    A = dense_matrix(nx)
    x = np.linspace(0, 1, nx)
    phi = np.cos(2 * np.pi * x) + x ** 2
    b = A @ phi
    got = o.poisson(b)
    assert np.max(np.abs(got - phi)) < 1e-9 * np.max(np.abs(phi))
    # converged to rounding: the extended-precision residual of the refined
    # solution is at the level of one rounding of each term
    Al = A.astype(np.longdouble)
    res = np.abs(b.astype(np.longdouble) - Al @ got.astype(np.longdouble)).astype(np.float64)
    scale = np.abs(A) @ np.abs(got)
    assert np.all(res <= 4 * EPS * scale)
    # while the unrefined band solve is visibly worse
    raw = o.poisson_approx(b)
    res_raw = np.abs(b.astype(np.longdouble) - Al @ raw.astype(np.longdouble)).astype(np.float64)
    assert res_raw.max() >= res.max()


# ---------------------------------------------------------------- moments
def test_charge_and_current_closed_forms_for_maxwell_juttner():
    """rho = q n (Sum_j f dp = integral of the normalised MJ) and, with A = 0,
    J = -(q^2/m) n <1/gamma> with <1/gamma> = K0(1/theta)/K1(1/theta) for the 1D
    Maxwell-Juttner distribution."""
    theta = 0.01
    p = landau(amplitude=0.0, theta=theta, np=512, pmin=-1.5, pmax=1.5, quadrature_depth=2)
    o = OracleRun(p)
    rho, J = o.moments()
    s = o.scalars()
    n_ref = p.n_ref
    q, m = s["charge"], s["mass"]
    assert np.allclose(rho, q * n_ref, rtol=1e-8)
    z = 1.0 / theta
    inv_gamma = sp.k0e(z) / sp.k1e(z)
    assert np.allclose(J, -(q * q / m) * n_ref * inv_gamma, rtol=1e-7)


def test_charge_follows_density_perturbation():
    p = landau(amplitude=0.05, k=4.0, quadrature_depth=3)
    o = OracleRun(p)
    rho, _ = o.moments()
    s = o.scalars()
    x = (np.arange(p.nx) + 0.5) * p.length / p.nx
    h = p.length / p.nx
    # cell average of 1 + a cos(kx)
    avg = 1 + p.amplitude * (np.sin(p.k * (x + h / 2)) - np.sin(p.k * (x - h / 2))) / (p.k * h)
    ref = s["charge"] * p.n_ref * avg
    assert np.allclose(rho, ref, rtol=1e-6)


# ---------------------------------------------------------------- time advance
def test_particle_number_conserved_with_walls():
    o = OracleRun(landau())
    dt = o.calc_dt()
    n0 = o.get("f1").sum()
    o.advance(dt, 30)
    n1 = o.get("f1").sum()
    assert abs(n1 / n0 - 1) < 1e-13


def test_stage_clock_accumulates_one_dt_per_step():
    o = OracleRun(landau(nx=16, np=32))
    dt = o.calc_dt()
    o.advance(dt, 7)
    assert o.scalars()["time"] == pytest.approx(7 * dt, rel=1e-14)


def test_fct_final_values_within_zalesak_bounds():
    """After one stage, every interior cell whose four faces are limited lies in
    [min, max] of {f^n, f^td} over its 5-point stencil (Zalesak 1979)."""
    o = OracleRun(two_stream(nx=64, np=128, length=4 * math.pi, amplitude=0.05))
    dt = o.calc_dt()
    o.advance(dt, 3)
    L = lib()
    L.vxo_stage_moments(o.h)
    L.vxo_stage_poisson(o.h)
    for sub in (0, 1, 2):
        L.vxo_stage_vlasov_sub(o.h, dt, 0, sub)
    f0 = o.get("f0", ghosts=1)
    f2 = o.get("f2", ghosts=1)
    f1 = o.get("f1", ghosts=1)
    lo = np.minimum(f0, f2)
    hi = np.maximum(f0, f2)

    def star(a, op):
        return op.reduce([a[1:-1, 1:-1], a[2:, 1:-1], a[:-2, 1:-1], a[1:-1, 2:], a[1:-1, :-2]])

    wmin = star(lo, np.minimum)
    wmax = star(hi, np.maximum)
    c = f1[1:-1, 1:-1]
    tol = 1e-12 * np.abs(f0).max()
    # columns i >= 1 and rows j >= 1 (the left column and bottom row sit on walls)
    sl = (slice(1, -1), slice(1, -1))
    assert np.all(c[sl] <= wmax[sl] + tol)
    assert np.all(c[sl] >= wmin[sl] - tol)


def test_uniform_in_x_fluxes_are_translation_invariant():
    """f(p) uniform in x, neutralised exactly (E = 0) and A = 0: the x-fluxes are
    the same on every face whose WENO stencil does not reach the zero ghosts,
    the p-fluxes vanish, and the first stage leaves those cells unchanged."""
    o = OracleRun(landau(amplitude=0.0, nx=32, np=64))
    dt = o.calc_dt()
    f_init = o.get("f1")
    L = lib()
    L.vxo_stage_moments(o.h)
    L.vxo_stage_poisson(o.h)
    assert np.abs(o.get("E")).max() == 0.0
    for sub in (0, 1, 2):
        L.vxo_stage_vlasov_sub(o.h, dt, 0, sub)
    FxH = o.get("FxH0")
    inner = FxH[3:30]
    assert np.all(inner == inner[0:1])
    assert np.abs(FxH[:, 1:-1]).max() > 0
    assert np.all(o.get("FpH0") == 0.0)
    f = o.get("f1")
    # equal fluxes cancel up to the rounding of (((f + a) + b) - b) - a
    assert np.allclose(f[4:-4], f_init[4:-4], rtol=8 * EPS, atol=0)


def test_mirror_symmetry_x_p_reflection():
    """Landau data with k L = 2 pi is invariant under (x, p) -> (L - x, -p).
    With E = 0 (exact neutralisation at t = 0) the stage-0 x-fluxes obey
    FxH(face i, row j) = -FxH(face N - i, row Np - 1 - j) on faces whose
    stencils stay off the walls, and the p-fluxes vanish."""
    p = landau(nx=64, np=128)
    o = OracleRun(p)
    dt = o.calc_dt()
    L = lib()
    L.vxo_stage_moments(o.h)
    L.vxo_stage_poisson(o.h)
    assert np.abs(o.get("E")).max() == 0.0
    L.vxo_stage_vlasov_sub(o.h, dt, 0, 0)
    G = o.get("FxH0", ghosts=1)  # row r holds face r - 1; column c holds p-row c - 1
    N = p.nx
    faces = np.arange(3, N - 2)
    a = G[faces + 1, 1:-1]
    b = -G[N - faces + 1, 1:-1][:, ::-1]
    F = G
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(F).max())
    assert np.abs(F).max() > 0
