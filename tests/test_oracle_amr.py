"""Pins of the adaptive-mesh part of the CPU oracle (oracle/oracle_amr.cpp).

The reference's own tests hold no vectors for the hierarchy, ghost fill,
coarse-fine fluxes or limiter synchronisation, and the reference cannot be
built here, so these are invariants the method fixes: disjoint patches on
each level, restriction equal to the mean of the finer cells, one limiter C
and one flux per shared face, and particle number conserved across levels,
steps and a regrid."""
import numpy as np
import pytest

from _amr_cases import particle_number, two_slab_laser
from _oracle import OracleRun


@pytest.fixture(scope="module")
def amr():
    o = OracleRun(two_slab_laser())
    yield o
    o.close()


def test_hierarchy_has_levels_and_disjoint_patches(amr):
    lay = amr.layout()
    assert set(lay[:, 0]) == {0, 1, 2}
    assert (lay[:, 0] == 0).sum() >= 4, "the test case must produce several finest patches"
    for lvl in (0, 1):
        rows = lay[lay[:, 0] == lvl]
        cover = {}
        for (_, x, p, nx, np_, *_rest) in rows:
            assert nx % 2 == 0 and np_ % 2 == 0 and x % 2 == 0 and p % 2 == 0
            for i in range(x, x + nx):
                for j in range(p, p + np_):
                    assert (i, j) not in cover
                    cover[(i, j)] = True


def _coarse_fine_pairs(lay):
    """(coarse level, coarse idx, fine idx) for patches one level apart that overlap."""
    out = []
    for lvl in (1, 2):
        c_rows = [(k, r) for k, r in enumerate(lay[lay[:, 0] == lvl])]
        f_rows = [(k, r) for k, r in enumerate(lay[lay[:, 0] == lvl - 1])]
        for kc, c in c_rows:
            for kf, f in f_rows:
                fx, fp, fnx, fnp = f[1] // 2, f[2] // 2, f[3] // 2, f[4] // 2
                if fx < c[1] + c[3] and fx + fnx > c[1] and fp < c[2] + c[4] and fp + fnp > c[2]:
                    out.append((lvl, kc, kf, c, f))
    return out


def test_restriction_is_mean_of_finer_cells():
    o = OracleRun(two_slab_laser())
    dt = o.calc_dt(0.4)
    o.advance(dt, 2)
    lay = o.layout()
    checked = 0
    for lvl, kc, kf, c, f in _coarse_fine_pairs(lay):
        cf = o.patch(lvl, kc, "f1")
        ff = o.patch(lvl - 1, kf, "f1")
        nest = o.patch(lvl, kc, "nested")
        for i in range(c[3]):
            for j in range(c[4]):
                if not nest[i, j]:
                    continue
                fi, fj = 2 * (c[1] + i) - f[1], 2 * (c[2] + j) - f[2]
                if not (0 <= fi < f[3] and 0 <= fj < f[4]):
                    continue  # covered by another finer patch
                t = 0.0
                for k in range(2):
                    for l in range(2):
                        t += ff[fi + k, fj + l]
                assert cf[i, j] == t / 4.0
                checked += 1
    assert checked > 100
    o.close()


def test_coarse_fluxes_are_means_of_finer_fluxes_and_c_is_shared():
    o = OracleRun(two_slab_laser())
    dt = o.calc_dt(0.4)
    o.advance(dt, 1)
    lay = o.layout()
    n_flux = n_c = 0
    for lvl, kc, kf, c, f in _coarse_fine_pairs(lay):
        cH = o.patch(lvl, kc, "FxH5", ghosts=1)
        fH = o.patch(lvl - 1, kf, "FxH5", ghosts=1)
        cC = o.patch(lvl, kc, "Cx", ghosts=1)
        fC = o.patch(lvl - 1, kf, "Cx", ghosts=1)
        # the finer patch's left and right edges, where they are interior to the coarse patch
        for fe in (f[1], f[1] + f[3]):
            ce = fe // 2 - c[1]
            if not (0 < ce < c[3]):
                continue
            for jf in range(0, f[4], 2):
                jc = (f[2] + jf) // 2 - c[2]
                if not (0 <= jc < c[4]):
                    continue
                a, b = fH[fe - f[1] + 1, jf + 1], fH[fe - f[1] + 1, jf + 2]
                assert cH[ce + 1, jc + 1] == (0.0 + a + b) * 0.25
                n_flux += 1
                assert fC[fe - f[1] + 1, jf + 1] == fC[fe - f[1] + 1, jf + 2] == cC[ce + 1, jc + 1]
                n_c += 1
    assert n_flux > 20 and n_c > 20
    o.close()


def test_particle_number_conserved_across_levels_steps_and_regrid():
    o = OracleRun(two_slab_laser())
    s = o.scalars()
    dx0, dp0 = s["dx"], s["dp"]  # coarsest cell sizes (base patch)
    dt = o.calc_dt(0.4)
    n0 = particle_number(o, dx0, dp0)
    o.advance(dt, 10)
    n1 = particle_number(o, dx0, dp0)
    lay_before = o.layout()
    o.regrid()
    n2 = particle_number(o, dx0, dp0)
    o.advance(dt, 5)
    n3 = particle_number(o, dx0, dp0)
    assert abs(n1 - n0) / n0 < 1e-13
    assert abs(n2 - n1) / n0 < 1e-13
    assert abs(n3 - n2) / n0 < 1e-13
    assert not np.array_equal(lay_before, o.layout()), "regrid should move the patches"
    o.close()


def _cubic_averages(c, lo, hi):
    """Exact average of c0 + c1 xi + c2 xi^2 + c3 xi^3 over [lo, hi]."""
    prim = lambda x: c[0] * x + c[1] * x ** 2 / 2 + c[2] * x ** 3 / 3 + c[3] * x ** 4 / 4
    return (prim(hi) - prim(lo)) / (hi - lo)


@pytest.mark.parametrize("which,levels", [(0, 2), (1, 1), (1, 3)])
def test_sub_cell_interpolants_are_exact_for_cubics(which, levels):
    """GetInterpolantsREF / REL (Rectangle.cpp:121-155, constants 76-110): from
    the averages of cells -2..2 of a cubic, the n sub-cells of cell 0 get the
    cubic's exact sub-cell averages (to the 15 digits of the interpolation
    table), and their mean is the cell's own value (the coarse cell is kept).
    A slipped table entry or a sub-cell bound off by one fails this."""
    p = two_slab_laser(levels=levels, nx=32, np=16)
    o = OracleRun(p)
    rng = np.random.default_rng(5 + which + levels)
    n_sub = 2 if which == 0 else 2 ** (levels - 1)
    for _ in range(20):
        c = rng.normal(size=4)
        f5 = [_cubic_averages(c, k - 0.5, k + 0.5) for k in range(-2, 3)]
        got = o.interp(which, f5)
        assert len(got) == n_sub
        want = [_cubic_averages(c, -0.5 + i / n_sub, -0.5 + (i + 1) / n_sub) for i in range(n_sub)]
        assert np.allclose(got, want, rtol=0, atol=1e-13 * (1 + np.abs(c).sum())), (got, want)
        assert abs(np.mean(got) - f5[2]) <= 1e-14 * (1 + np.abs(c).sum())
    o.close()
