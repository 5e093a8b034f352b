"""vx_patch_upload / vx_patch_download: an f slot of a patch to and from a
host array with 0, 1 or 2 ghost layers (the pitched plane on the device,
row-major on the host).  Checked by round trips: what goes up comes back bit
for bit at every ghost width, the layers outside the copied window keep their
contents, and a run whose state went down and up again steps exactly as one
that never left the device."""
import math
import sys

import numpy as np
import pytest

import veritas_amd as va
from veritas_amd import params as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nx,np_", [(64, 128), (130, 60), (8, 38)])
def test_round_trip_every_ghost_width(nx, np_):
    say(f"case: patch io round trip {nx}x{np_}")
    r = va.Run(P.landau(nx=nx, np=np_))
    rng = np.random.default_rng(nx * 1000 + np_)
    full = rng.standard_normal((nx + 4, np_ + 4))
    r.put("f2", full, ghosts=2)
    for g in (0, 1, 2):
        got = r.get("f2", ghosts=g)
        assert np.array_equal(got, full[2 - g:nx + 2 + g, 2 - g:np_ + 2 + g]), g
    inner = rng.standard_normal((nx + 2, np_ + 2))
    r.put("f2", inner, ghosts=1)  # the outer ghost layer stays
    want = full.copy()
    want[1:nx + 3, 1:np_ + 3] = inner
    assert np.array_equal(r.get("f2", ghosts=2), want)
    with pytest.raises(KeyError):
        r.put("f2", inner, ghosts=2)  # wrong size for two layers
    r.close()


def test_state_that_left_the_device_steps_the_same():
    say("case: two-stream 128x256, f0/f1 down and up before stepping")
    p = P.two_stream(nx=128, np=256, length=8 * math.pi, amplitude=0.05)
    a, b = va.Run(p), va.Run(p)
    for s in ("f0", "f1"):
        b.put(s, b.get(s, ghosts=2), ghosts=2)
    for _ in range(5):
        a.advance(a.calc_dt(0.4))
        b.advance(b.calc_dt(0.4))
    assert np.array_equal(a.get("f1"), b.get("f1"))
    a.close()
    b.close()
