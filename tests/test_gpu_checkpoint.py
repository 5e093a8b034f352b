"""Checkpoint/restart: a run restored from a checkpoint continues bit for bit
as the run that wrote it (single walled patch on the fused path, and the
three-level hierarchy after a regrid)."""
import numpy as np
import pytest

import veritas_amd as va
from _amr_cases import two_slab_laser
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def same_state(a, b, tag):
    la, lb = a.layout(), b.layout()
    assert np.array_equal(la, lb), tag
    for lvl in sorted(set(la[:, 0])):
        for k in range(int((la[:, 0] == lvl).sum())):
            assert_same(f"{tag} level {lvl} patch {k}", a.patch(lvl, k, "f1", ghosts=2),
                        b.patch(lvl, k, "f1", ghosts=2))
    for w in ("phi", "E", "By", "Bz", "Ay", "Az", "a2", "scalars"):
        assert_same(f"{tag} {w}", a.get(w), b.get(w))


def test_restart_single_patch_fused(tmp_path):
    p = va.two_stream(nx=128, np=128, length=8 * np.pi, amplitude=0.02)
    a = va.Run(p, fused=True)
    dt = a.calc_dt(0.4)
    a.advance(dt, 5)
    a.save(tmp_path / "ck.bin")
    a.advance(dt, 5)
    b = va.Run(p, fused=True)
    b.load(tmp_path / "ck.bin")
    b.advance(dt, 5)
    same_state(a, b, "fused restart")
    a.close()
    b.close()


def test_restart_three_levels_after_regrid(tmp_path):
    p = two_slab_laser()
    a = va.Run(p, fused=False)
    dt = a.calc_dt(0.4)
    a.advance(dt, 2)
    a.regrid()
    a.advance(dt, 1)
    a.save(tmp_path / "ck.bin")
    a.advance(dt, 3)
    b = va.Run(p, fused=False)
    b.load(tmp_path / "ck.bin")
    b.advance(dt, 3)
    same_state(a, b, "AMR restart")
    a.close()
    b.close()
