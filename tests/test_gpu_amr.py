"""GPU parity of the adaptive mesh and of the laser drive: hierarchy layout,
refinement flags, per-patch f after steps with coarse-fine ghosts, coarse-fine
fluxes, limiter-C sync and restriction, and the regrid transfer, all against
the CPU oracle bit for bit; transverse fields and A^2 are non-zero here."""
import sys

import numpy as np
import pytest

import veritas_amd as va
from _amr_cases import particle_number, two_slab_laser
from _oracle import OracleRun
from test_gpu_parity import assert_same, assert_ulp

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare_mesh(g, o, tag, what=("f1", "f0")):
    lay_g, lay_o = g.layout(), o.layout()
    assert np.array_equal(lay_g, lay_o), f"{tag}: layouts differ\n{lay_g}\n{lay_o}"
    for lvl in sorted(set(lay_o[:, 0])):
        n = int((lay_o[:, 0] == lvl).sum())
        for k in range(n):
            for w in what:
                assert_same(f"{tag} level {lvl} patch {k} {w}", g.patch(lvl, k, w, ghosts=2),
                            o.patch(lvl, k, w, ghosts=2))


def compare_fields(g, o, tag):
    assert_ulp(f"{tag} phi", g.get("phi"), o.get("phi"))
    assert_ulp(f"{tag} E", g.get("E"), o.get("E"))
    for w in ("By", "Bz", "Ey", "Ez", "Ay", "Az", "a2"):
        assert_same(f"{tag} {w}", g.get(w), o.get(w))


def test_laser_single_level_fused_matches():
    """One walled patch: the fused stage kernel with a non-zero A^2 and the
    laser-driven Maxwell stage."""
    p = two_slab_laser(levels=1)
    say("case: laser single level fused, 12 steps")
    g, o = va.Run(p, fused=True), OracleRun(p)
    dt = o.calc_dt(0.4)
    assert g.calc_dt(0.4) == dt
    g.advance(dt, 12)
    o.advance(dt, 12)
    assert_same("f1", g.get("f1"), o.get("f1"))
    compare_fields(g, o, "laser")
    assert np.abs(o.get("Bz")).max() > 0 and np.abs(o.get("a2")).max() > 0
    g.close()
    o.close()


@pytest.fixture(scope="module")
def amr_pair():
    p = two_slab_laser()
    say("case: two-slab laser, 3 levels")
    g, o = va.Run(p, fused=False), OracleRun(p)
    yield p, g, o
    g.close()
    o.close()


def test_amr_initial_hierarchy_matches(amr_pair):
    p, g, o = amr_pair
    say("step: initial hierarchy")
    compare_mesh(g, o, "initial")
    for lvl in range(2):
        assert np.array_equal(g.flags(lvl), o.flags(lvl)), f"flags of hierarchy level {lvl}"


def test_device_flag_signatures(amr_pair):
    """The flags come back compacted on the device (in the reference's scan
    order: equal to the oracle's list above); the row and column signatures
    computed there equal the counts of that list over each patch's range."""
    p, g, o = amr_pair
    lay = g.layout()
    nlev = int(lay[:, 0].max()) + 1
    checked = 0
    for lvl in range(nlev - 1):  # levels that flag cells for a finer one (0 = coarsest)
        fl = g.flags(lvl)
        assert np.array_equal(fl, o.flags(lvl))
        npatch = int((lay[:, 0] == nlev - 1 - lvl).sum())
        for k in range(npatch):
            sg = g.signatures(lvl, k)
            x0, p0, rows, cols = sg["lo_x"], sg["lo_p"], sg["rows"], sg["cols"]
            inside = ((fl[:, 0] >= x0) & (fl[:, 0] < x0 + len(rows)) & (fl[:, 1] >= p0) &
                      (fl[:, 1] < p0 + len(cols)))
            sub = fl[inside]
            assert np.array_equal(rows, np.bincount(sub[:, 0] - x0, minlength=len(rows)))
            assert np.array_equal(cols, np.bincount(sub[:, 1] - p0, minlength=len(cols)))
            checked += int(rows.sum())
    assert checked > 0


def test_amr_steps_and_regrid_match(amr_pair):
    p, g, o = amr_pair
    dt = o.calc_dt(0.4)
    assert g.calc_dt(0.4) == dt
    for n, tag in ((1, "after 1 step"), (3, "after 4 steps")):
        say(f"step: advance {n}")
        g.advance(dt, n)
        o.advance(dt, n)
        compare_mesh(g, o, tag)
        compare_fields(g, o, tag)
    for lvl in range(2):
        assert np.array_equal(g.flags(lvl), o.flags(lvl)), f"flags of hierarchy level {lvl}"
    say("step: regrid")
    g.regrid()
    o.regrid()
    n, b = g.regrid_stats()
    say(f"  regrids {n}, rebuilt {b}")
    assert b >= 1, "this regrid is expected to change the boxes (it exercises the transfer)"
    compare_mesh(g, o, "after regrid")
    say("step: advance 3 after regrid")
    g.advance(dt, 3)
    o.advance(dt, 3)
    compare_mesh(g, o, "after regrid + 3 steps")
    compare_fields(g, o, "after regrid + 3 steps")
    s = o.scalars()
    n_g = particle_number(g_as_oracle_layout(g, o), s["dx"], s["dp"])
    n_o = particle_number(o, s["dx"], s["dp"])
    assert n_g == n_o


def test_regrid_to_the_same_hierarchy_keeps_the_patches(amr_pair):
    """A regrid right after a regrid sees the same data and finds the same
    boxes: the GPU run keeps its patches while the oracle rebuilds them with the
    reference's transfer; the states agree then and three steps later."""
    p, g, o = amr_pair
    dt = o.calc_dt(0.4)
    kept = False
    for k in range(3):  # regrid (both runs) until the boxes come out unchanged
        say(f"step: regrid {k}")
        n0, b0 = g.regrid_stats()
        lay = g.layout().copy()
        g.regrid()
        o.regrid()
        n1, b1 = g.regrid_stats()
        assert n1 == n0 + 1
        compare_mesh(g, o, f"after regrid {k}")
        if b1 == b0:
            kept = True
            assert np.array_equal(g.layout(), lay)
            break
    assert kept, "no regrid kept the hierarchy"
    compare_mesh(g, o, "after an unchanged regrid")
    g.advance(dt, 3)
    o.advance(dt, 3)
    compare_mesh(g, o, "after an unchanged regrid + 3 steps")
    compare_fields(g, o, "after an unchanged regrid + 3 steps")


@pytest.mark.parametrize("np_", [64, 256])
def test_moment_forms_agree_on_every_level(np_):
    """The p-moments in every form (the split form of large launches, the
    producer/consumer pairs of small ones and the wide form of relativeToBottom
    1 patches, pinned with VX_MOMENTS_FORM) on the three-level case, whose
    patches have relativeToBottom 1, 2 and 4, and whose columns have the
    laser's A^2 (read from a logarithm table or formed per column): the same
    bits, and the oracle's.  With 256 base p-cells every level's columns span
    several 64-cell chunks, whose logarithm terms carry the last two of the
    chunk before."""
    import os
    p = two_slab_laser(np=np_)
    o = OracleRun(p)
    dt = o.calc_dt(0.4)
    runs = {}
    for form in ("split", "pair", "wide"):
        say(f"case: moments form {form}, 3 levels, np {np_}, 3 steps")
        os.environ["VX_MOMENTS_FORM"] = form
        try:
            g = va.Run(p, fused=False)
            g.advance(dt, 3)
            runs[form] = g
        finally:
            del os.environ["VX_MOMENTS_FORM"]
    o.advance(dt, 3)
    compare_mesh(runs["split"], o, "split form")
    compare_mesh(runs["pair"], o, "pair form")
    compare_mesh(runs["wide"], o, "wide form")
    compare_fields(runs["split"], o, "split form")
    for g in runs.values():
        g.close()
    o.close()


class g_as_oracle_layout:
    """GPU f1 with the oracle's nested masks (the layouts are equal)."""

    def __init__(self, g, o):
        self.g, self.o = g, o

    def layout(self):
        return self.g.layout()

    def patch(self, lvl, k, w):
        return self.o.patch(lvl, k, w) if w == "nested" else self.g.patch(lvl, k, w)


@pytest.mark.parametrize("policy", ["8,4", "128,64", "-512,-1024"])
def test_batched_strip_policies_change_no_bits(policy, monkeypatch):
    """The strip policy of the batched sub-step-0 flux and limiter launches
    (VX_AMR_POLICY pins it; the autotuner picks one per device and set of
    patch shapes) changes no bit: three steps of the three-level case equal to
    the oracle with short, tall and wavefront-count strips."""
    monkeypatch.setenv("VX_AMR_POLICY", policy)
    p = two_slab_laser()
    say(f"case: three-level two-slab laser, strip policy {policy}, 3 steps vs oracle")
    g, o = va.Run(p, fused=False), OracleRun(p)
    dt = o.calc_dt(0.4)
    for _ in range(3):
        g.advance(dt)
        o.advance(dt)
    compare_mesh(g, o, f"policy {policy}")
    compare_fields(g, o, f"policy {policy}")
    g.close()
    o.close()


def test_per_patch_launches_equal_the_batched_launches():
    """The plain per-patch form (every sub-step launched patch by patch) and
    the batched form (one launch per sub-step for every patch) give the same
    bits on the three-level case, three steps from the same state."""
    p = two_slab_laser()
    say("case: three-level two-slab laser, per-patch vs batched launches, 3 steps")
    a, b = va.Run(p, fused=False), va.Run(p, fused=False, batched=False)
    dt = a.calc_dt(0.4)
    for _ in range(3):
        a.advance(dt)
        b.advance(dt)
    compare_mesh(a, b, "per-patch vs batched")
    compare_fields(a, b, "per-patch vs batched")
    a.close()
    b.close()


def test_refinement_ratio_four():
    """Refinement ratio 4 (the task programs are compiled per ratio): the
    initial two-level hierarchy of 7 finer patches (ghost fill from the
    coarser level and restriction at r = 4) bitwise against the oracle.  A step
    is refused with an error: the reference keeps 2 ghost cells per patch
    (Rectangle.hpp:96-97), and for r = 4 its coarse-fine fluxes
    (CalculateFluxToCoarseX/P, Rectangle.cpp:1313-1394) read finer fluxes up to
    4 cells beyond the finer patch, and its coarse-to-fine transfer
    (Rectangle.cpp:892-918) writes that far, outside the patch's data; the
    device checks every task against the patch bounds first."""
    p = two_slab_laser(r=4, levels=2)
    say("case: two-slab laser, 2 levels, ratio 4")
    g, o = va.Run(p, fused=False), OracleRun(p)
    compare_mesh(g, o, "r=4 initial")
    assert int(g.layout()[:, 0].max()) == 1 and len(g.layout()) > 2
    dt = o.calc_dt(0.4)
    assert g.calc_dt(0.4) == dt
    say("step: advance 1 (refused)")
    with pytest.raises(va.VeritasError, match="out of range"):
        g.advance(dt, 1)
    g.close()
    o.close()


@pytest.mark.parametrize("data", ["bilinear", "gaussian"])
def test_fills_of_two_touching_patches(data):
    """The case of the oracle's fill pins (test_oracle_amr_fill.py): two
    touching finer patches placed by RefinementOverride, whose ghost cells come
    from the coarser level (edges and corner blocks) and from each other, then
    a regrid to a shifted region (the new patches filled from the old ones and
    from the coarser level).  On the device bitwise against the oracle; with
    bilinear data every finer cell and ghost equals the exact fine-cell
    average, with Gaussian data every interpolated block averages to its coarse
    cell."""
    import test_oracle_amr_fill as F
    f = getattr(F, data)
    say(f"case: two touching finer patches, {data} data")
    g, gflag = F.make_run(f, F.L_SHAPE, factory=lambda d, h: va.Run.from_desc(d, h, fused=False))
    o, oflag = F.make_run(f, F.L_SHAPE)
    compare_mesh(g, o, f"{data} initial")
    assert len(F.fine_patches(g)) == 2
    old = F.boxes_of(g)
    if data == "bilinear":
        F.check_bilinear(g, "device, initial")
    else:
        F.check_block_means(g, "device, initial")
    say("step: regrid to the shifted region")
    gflag["boxes"] = oflag["boxes"] = F.SHIFTED
    g.regrid()
    o.regrid()
    assert F.boxes_of(g) != old
    compare_mesh(g, o, f"{data} after the regrid")
    if data == "bilinear":
        F.check_bilinear(g, "device, after the regrid")
    else:
        F.check_block_means(g, "device, after the regrid", interior_from=old)
    g.close()
    o.close()
