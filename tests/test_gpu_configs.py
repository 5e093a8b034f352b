"""The benchmark configurations at their full size on the GPU.

- The laser configuration (base grid 8192 x 512, three levels at ratio 2,
  flagged where f > 1e-6 max f through the reference's size weight): the
  initial hierarchy and the refinement flags of every level equal the
  oracle's, then two steps, a regrid and one more step bit for bit.
- The serving stream: the 256 seeded scan requests (Landau, two-stream and
  laser runs of 1-3 levels, Nx 128-4096, Np 256-2048) replayed with Poisson
  arrivals; every request's result is bit for bit the result it gives alone."""
import sys
import time

import numpy as np
import pytest

import veritas_amd as va
from _oracle import OracleRun
from test_gpu_amr import compare_fields, compare_mesh

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_laser_configuration_at_full_size():
    p = va.laser()
    say("case: laser base 8192x512, 3 levels: set-up, flags, 2 steps, regrid, 1 step")
    t0 = time.time()
    o = OracleRun(p)
    say(f"  oracle set up ({time.time() - t0:.0f} s)")
    g = va.Run(p)
    say(f"  GPU run set up ({time.time() - t0:.0f} s)")
    compare_mesh(g, o, "set-up", what=("f1",))
    for lvl in range(3):
        assert np.array_equal(g.flags(lvl), o.flags(lvl)), f"flags of level {lvl}"
    say(f"  layouts and flags equal ({time.time() - t0:.0f} s)")
    dt = o.calc_dt(0.4)
    assert g.calc_dt(0.4) == dt
    g.advance(dt, 2)
    o.advance(dt, 2)
    compare_mesh(g, o, "2 steps")
    compare_fields(g, o, "2 steps")
    say(f"  2 steps equal ({time.time() - t0:.0f} s)")
    g.regrid(2 * dt)
    o.regrid()
    compare_mesh(g, o, "regrid", what=("f1",))
    g.advance(dt, 1)
    o.advance(dt, 1)
    compare_mesh(g, o, "regrid + 1 step")
    compare_fields(g, o, "regrid + 1 step")
    say(f"  regrid and one more step equal ({time.time() - t0:.0f} s)")
    g.close()
    o.close()


def test_scan_stream_results_are_their_solo_runs():
    from veritas_amd.serve import poisson_arrivals, scan_requests
    reqs = scan_requests(256, seed=5)
    # about 80 % of the capacity measured for this stream on an MI355X
    arrivals = poisson_arrivals(len(reqs), 10.0, seed=7)
    say(f"case: serving {len(reqs)} scan requests with Poisson arrivals, keep state")
    t0 = time.time()
    srv = va.Server(budget_bytes=32 << 30, max_active=16, keep_state=True)
    for k, (r, t) in enumerate(zip(reqs, arrivals)):
        srv.submit(r["params"], r["steps"], arrival=float(t), rid=k)
    srv.run()
    rec = {int(r["id"]): r for r in srv.records()}
    assert len(rec) == len(reqs)
    st = srv.stats()
    say(f"  served in {time.time() - t0:.0f} s: {st}")
    assert st["peak_active"] >= 2 and st["batched_launches"] > 0
    for k, r in enumerate(reqs):
        run = va.Run(r["params"])
        for _ in range(r["steps"]):
            run.advance(run.calc_dt(0.4))
        f = run.get("f1").ravel()
        run.close()
        g = srv.final_f(k)
        assert np.array_equal(g, f), f"request {k} ({r['kind']}, {r['params'].nx}x{r['params'].np})"
        if k % 32 == 31:
            say(f"  {k + 1} requests equal to their solo runs ({time.time() - t0:.0f} s)")
    srv.close()
