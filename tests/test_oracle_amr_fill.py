"""Pins of the oracle's coarse-to-fine fills by what the mathematics fixes.

A two-level hierarchy (ratio 2) whose finer patches are placed by the
RefinementOverride hook: an L-shaped flagged region that the clustering cuts
into two touching patches, so that each finer patch has ghost cells filled
from the coarser level (edges and corner blocks, GetWenoValueFromCoarseLevel,
Rectangle.cpp:343-415, 417-559) and from the other patch (same-level copies,
Rectangle.cpp:564-614).  A regrid to a shifted region then fills the new
patches from the old ones and, where they did not reach, from the coarser
level (GetDataFromCoarseLevelRectangle, Rectangle.cpp:892-918).

- The fill is the tensor product of the 1-D REF interpolants (exact for
  cubics, Rectangle.cpp:121-137) with a mean correction.  Cell values start as
  midpoint-rule averages (Rectangle.cpp:620-662), exact for a bilinear f: so
  every ghost cell and every filled cell must equal the exact average of the
  bilinear f over that fine cell, to rounding.  A transposed sub-cell index, a
  wrong ghost layer or a wrong neighbour breaks this.
- The fill conserves each coarse cell: the 1-D REF interpolants' coefficient
  vectors sum to zero over the r sub-cells and the correction restores the
  rounding, so for a Gaussian f (not reproduced exactly) each r x r block of
  fine cells interpolated from one coarse cell averages to that cell's value,
  to rounding.  A block filled from the wrong coarse cell or mixing two
  breaks this."""
import math

import numpy as np

from veritas_amd import params as P
from _oracle import OracleRun

MC = P.M_E * P.C
NX = NP = 64
DX = 1.0e-7
PMIN, DP = -0.5 * MC, MC / NP
LX, LP = NX * DX, NP * DP
R = 2


def bilinear(x, p):
    return (1.0 + 0.7 * x / LX) * (2.0 + 0.9 * (p - PMIN) / LP)


def gaussian(x, p):
    u, v = (x / LX - 0.55) / 0.08, ((p - PMIN) / LP - 0.6) / 0.07
    return 1.0 + 3.0 * math.exp(-0.5 * (u * u + v * v))


def make_run(f, boxes, factory=OracleRun.from_desc):
    """A run (of the oracle, or of `factory(desc, hooks)`) from f whose base
    cells (i, j) inside any of `boxes` (base-cell index ranges) are flagged; the
    base is depth 1, the finer level depth 0.  Returns the run and the dict
    holding the boxes (change them, then regrid)."""
    flagged = {"boxes": boxes}

    def override(x, p, depth, sp):
        if depth != 1:
            return False
        i, j = x / DX, (p - PMIN) / DP
        return any(x0 <= i < x1 and p0 <= j < p1 for (x0, x1), (p0, p1) in flagged["boxes"])

    d = P.SettingsDesc.make(nx=NX, dx=DX, Lfinest=2, r=R, minEfficiency=0.9, refinementCriteria=1e300,
                            quadratureDepth=1,
                            species=[dict(mass=P.M_E, charge=-P.Q_E, np=NP, dp=DP, pmin=PMIN)])
    hooks = P.Hooks(lambda x, p, sp: f(x, p), refinement_override=override)
    run = factory(d, hooks)
    run._hooks = hooks  # the callbacks live as long as the run
    return run, flagged


L_SHAPE = [((20, 36), (24, 40)), ((28, 36), (40, 48))]
SHIFTED = [((22, 38), (24, 40)), ((30, 38), (40, 48))]


def fine_patches(run):
    lay = run.layout()
    return [(k, row) for k, row in enumerate(lay[lay[:, 0] == 0])]


def fine_centres(row, g=2):
    xp, pp, n_x, n_p = (int(v) for v in row[1:5])
    ii = np.arange(-g, n_x + g)[:, None]
    jj = np.arange(-g, n_p + g)[None, :]
    return (xp + ii + 0.5) * DX / R, PMIN + (pp + jj + 0.5) * DP / R


def check_bilinear(run, tag):
    for k, row in fine_patches(run):
        got = run.patch(0, k, "f1", ghosts=2)
        x, p = fine_centres(row)
        want = bilinear(x, p)
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-13, f"{tag}: finer patch {k} {row[1:5].tolist()} off by {err:.3g}"


def check_block_means(run, tag, interior_from=None):
    """Every r x r block of ghost cells (and, with `interior_from`, of interior
    cells not covered by those earlier patches) whose cells lie in no finer
    patch was interpolated from one coarse cell: its mean is that cell."""
    coarse = run.patch(1, 0, "f1")
    boxes = boxes_of(run)
    checked = 0
    for k, row in fine_patches(run):
        xp, pp, n_x, n_p = (int(v) for v in row[1:5])
        got = run.patch(0, k, "f1", ghosts=2)
        for bi in range(-2, n_x + 2, R):
            for bj in range(-2, n_p + 2, R):
                gi, gj = xp + bi, pp + bj
                ghost = not (0 <= bi < n_x and 0 <= bj < n_p)
                if ghost:
                    if any(covered_by(boxes, gi + a, gj + b) for a in range(R) for b in range(R)):
                        continue
                elif interior_from is None or covered_by(interior_from, gi, gj):
                    continue
                blk = got[bi + 2:bi + 2 + R, bj + 2:bj + 2 + R]
                parent = coarse[gi // R, gj // R]
                assert abs(blk.mean() - parent) <= 1e-14 * abs(parent), (
                    f"{tag}: finer patch {k} block ({bi}, {bj}) mean {blk.mean()!r} vs coarse {parent!r}")
                checked += 1
    assert checked > 0
    return checked


def boxes_of(run):
    return [tuple(int(v) for v in row[1:5]) for _, row in fine_patches(run)]


def covered_by(boxes, gi, gj):
    """Whether fine cell (gi, gj) (finer-level indices) lies in one of the boxes."""
    return any(x0 <= gi < x0 + nx_ and p0 <= gj < p0 + np_ for x0, p0, nx_, np_ in boxes)


def test_ghost_fill_is_exact_for_bilinear_data():
    run, _ = make_run(bilinear, L_SHAPE)
    pts = fine_patches(run)
    assert len(pts) == 2, run.layout()
    # the two patches touch along p: ghost cells of each lie in the other
    (_, a), (_, b) = pts
    assert a[1] < b[1] + b[3] and b[1] < a[1] + a[3] and (a[2] + a[4] == b[2] or b[2] + b[4] == a[2])
    check_bilinear(run, "initial hierarchy")
    run.close()


def test_regrid_transfer_is_exact_for_bilinear_data():
    run, flagged = make_run(bilinear, L_SHAPE)
    old = boxes_of(run)
    flagged["boxes"] = SHIFTED
    run.regrid()
    new = boxes_of(run)
    assert len(new) == 2 and new != old
    # part of each new patch lies outside the old ones: filled from the coarse level
    assert any(not all(covered_by(old, xp + i, pp + j) for i in range(n_x) for j in range(n_p))
               for xp, pp, n_x, n_p in new)
    check_bilinear(run, "after the regrid")
    run.close()


def test_interpolated_blocks_average_to_their_coarse_cell():
    run, flagged = make_run(gaussian, L_SHAPE)
    n0 = check_block_means(run, "initial ghosts")
    old = boxes_of(run)
    flagged["boxes"] = SHIFTED
    run.regrid()
    n1 = check_block_means(run, "after the regrid", interior_from=old)
    assert n0 > 20 and n1 > n0
    run.close()
