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


@pytest.mark.parametrize("mixed", [False, True])
def test_restart_single_patch_fused(tmp_path, mixed):
    p = va.two_stream(nx=128, np=128, length=8 * np.pi, amplitude=0.02)
    a = va.Run(p, fused=True, mixed=mixed)
    dt = a.calc_dt(0.4)
    a.advance(dt, 5)
    a.save(tmp_path / "ck.bin")
    a.advance(dt, 5)
    b = va.Run(p, fused=True, mixed=mixed)
    b.load(tmp_path / "ck.bin")
    b.advance(dt, 5)
    same_state(a, b, "fused restart" + (" (mixed)" if mixed else ""))
    a.close()
    b.close()


def test_checkpoint_of_other_inputs_or_truncated_is_rejected(tmp_path):
    """A file written by a run of another scenario on the same grid, or with
    the other precision, is refused; a truncated file is refused before the
    loading run changes (it then steps as if nothing happened)."""
    p = va.two_stream(nx=64, np=64, length=8 * np.pi, amplitude=0.02)
    a = va.Run(p)
    dt = a.calc_dt(0.4)
    a.advance(dt, 2)
    a.save(tmp_path / "ck.bin")
    other = va.Run(p.replace(drift=0.25))
    with pytest.raises(va.VeritasError, match="other inputs"):
        other.load(tmp_path / "ck.bin")
    mixed = va.Run(p, mixed=True)
    with pytest.raises(va.VeritasError, match="other inputs"):
        mixed.load(tmp_path / "ck.bin")
    data = (tmp_path / "ck.bin").read_bytes()
    (tmp_path / "cut.bin").write_bytes(data[: len(data) - 1000])
    b, c = va.Run(p), va.Run(p)
    b.advance(dt, 1)
    c.advance(dt, 1)
    with pytest.raises(va.VeritasError, match="truncated"):
        b.load(tmp_path / "cut.bin")
    b.advance(dt, 2)
    c.advance(dt, 2)
    same_state(b, c, "after a refused load")
    for r in (a, other, mixed, b, c):
        r.close()


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
