"""The serving loop (host/serve.cpp): runs admitted by device-memory footprint
while others are on the GPU, stepped together with their single walled
patches in shared stage launches.  Every result must be the one the request
gives alone (bit for bit: the same kernels with the same arguments, and dt =
0.4 x CFL before every step in both), the budget and the slot limit must
hold, and requests that arrive during a run must be admitted before it ends."""
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


def alone(p, steps, mixed=False):
    r = va.Run(p, mixed=mixed)
    for _ in range(steps):
        r.advance(r.calc_dt(0.4))
    f = r.get("f1").ravel()
    t = r.scalars()["time"]
    r.close()
    return f, t


CASES = [
    ("landau 128x256", P.landau(nx=128, np=256), 30, 0.0, False),
    ("two-stream 256x512", P.two_stream(nx=256, np=512, length=8 * math.pi, amplitude=0.05), 25, 0.0, False),
    ("laser 1 level 128x256", P.laser(nx=128, np=256, levels=1, wavelengths=8, ramp=2), 30, 0.02, False),
    ("landau 256x256 mixed", P.landau(nx=256, np=256, k=3.0, length=2 * math.pi / 3), 30, 0.04, True),
    ("laser 2 levels 128x64", P.laser(nx=128, np=64, levels=2, wavelengths=40, ramp=5, gap=8, pmin=-0.5,
                                     pmax=0.5, theta=0.01, refinement_criteria=0.01, min_efficiency=0.85,
                                     size_weight=0.0),
     12, 0.06, False),
    ("two-stream 512x256", P.two_stream(nx=512, np=256, length=8 * math.pi, amplitude=0.05), 20, 0.1, False),
]


def test_each_result_is_the_same_alone_and_in_the_mix():
    say("case: serve 6 requests with staggered arrivals, keep state")
    srv = va.Server(budget_bytes=32 << 30, max_active=8, keep_state=True)
    for k, (name, p, steps, arr, mixed) in enumerate(CASES):
        srv.submit(p, steps, arrival=arr, mixed=mixed, rid=k)
    srv.run()
    rec = {int(r["id"]): r for r in srv.records()}
    st = srv.stats()
    say(f"  stats {st}")
    assert len(rec) == len(CASES)
    assert st["batched_launches"] > 0 and st["peak_active"] >= 2
    for k, (name, p, steps, arr, mixed) in enumerate(CASES):
        say(f"  alone: {name}")
        f, t = alone(p, steps, mixed)
        g = srv.final_f(k)
        assert g.shape == f.shape, name
        assert np.array_equal(g, f), f"{name}: {np.count_nonzero(g != f)} entries differ"
        assert rec[k]["time"] == t, name
        assert rec[k]["finished"] >= rec[k]["started"] >= rec[k]["admitted"] >= rec[k]["arrival"] - 1e-9
    # the single walled patches shared launches; the two-level run did not
    assert rec[0]["batched"] == 1.0 and rec[4]["batched"] == 0.0
    srv.close()


def steps_lasting(p, seconds):
    """Steps of scenario p that take about `seconds` alone on this GPU (from a
    timed run of 5 steps), so that the timing assertions below hold whatever
    the speed of the host and the device."""
    import time
    r = va.Run(p)
    dt = r.calc_dt(0.4)
    r.advance(dt, 2)
    r.sync()
    t0 = time.perf_counter()
    r.advance(dt, 5)
    r.sync()
    per = (time.perf_counter() - t0) / 5
    r.close()
    return max(20, int(seconds / per))


def test_admission_by_footprint_and_slots():
    p = P.two_stream(nx=512, np=512, length=8 * math.pi, amplitude=0.05)
    say("case: serve equal requests under a budget below one run, then of 2.5 runs")
    srv = va.Server(budget_bytes=1, max_active=8)  # below one run: one at a time
    for k in range(3):
        srv.submit(p, 10, arrival=0.0, rid=k)
    srv.run()
    rec = sorted(srv.records(), key=lambda r: r["admitted"])
    st = srv.stats()
    assert st["peak_active"] == 1
    for a, b in zip(rec, rec[1:]):
        assert b["admitted"] >= a["finished"] - 1e-6  # each waited for memory
    fp = rec[0]["footprint"]
    srv.close()
    # runs long enough (about a second each) that the next one is built while one runs
    steps = steps_lasting(p, 1.0)
    srv = va.Server(budget_bytes=int(2.5 * fp), max_active=3)
    for k in range(4):
        srv.submit(p, steps, arrival=0.0, rid=k)
    srv.run()
    st = srv.stats()
    say(f"  footprint {fp:.3e} B, stats {st}")
    assert st["peak_active"] == 2 and st["peak_bytes"] <= 2.5 * fp
    srv.close()


def test_arrivals_during_a_run_are_admitted_before_it_ends():
    say("case: one long run, short ones arriving while it runs")
    long_p = P.two_stream(nx=1024, np=1024, length=8 * math.pi, amplitude=0.05)
    steps = steps_lasting(long_p, 2.0)  # the long run steps for about 2 s after its set-up
    srv = va.Server(budget_bytes=32 << 30, max_active=8)
    srv.submit(long_p, steps, arrival=0.0, rid=0)
    for k, arr in enumerate((0.05, 0.1, 0.15)):
        srv.submit(P.landau(nx=128, np=256), 20, arrival=arr, rid=k + 1)
    srv.run()
    rec = {int(r["id"]): r for r in srv.records()}
    for k in (1, 2, 3):
        assert rec[k]["admitted"] < rec[0]["finished"]
        assert rec[k]["finished"] < rec[0]["finished"]
    srv.close()
