"""Steps replayed as recorded hipGraphs (SolverManager::Advance,
SolverManager.cpp:28-39) against the CPU oracle, bit for bit, after every
replay.  Between replays the test downloads state (host synchronisation and
eager copies), so every replay is checked after other work on the stream.

Cases: the fused single-patch path, whose three f planes cycle with period
two (two graphs); the laser drive, whose hook values reach the recorded
Maxwell stages through device memory; the three-level adaptive mesh, across a
regrid that discards the recorded steps."""
import math
import sys

import numpy as np
import pytest

import veritas_amd as va
from _amr_cases import two_slab_laser
from _oracle import OracleRun
from test_gpu_amr import compare_fields, compare_mesh
from test_gpu_parity import assert_same
from veritas_amd.params import landau, two_stream

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_fused_replays_match_every_step():
    p = two_stream(nx=64, np=128, length=4 * math.pi, amplitude=0.05)
    say("case: graphs, fused two-stream 64x128, 10 steps one at a time")
    g, o = va.Run(p, fused=True, graphs=True), OracleRun(p)
    dt = o.calc_dt()
    for step in range(10):
        g.advance(dt)
        o.advance(dt)
        assert_same(f"step {step} f1", g.get("f1", 2), o.get("f1", 2))
        assert_same(f"step {step} E", g.get("E"), o.get("E"))
        assert g.scalars()["time"] == o.scalars()["time"]
    st = g.graph_stats()
    assert st["records"] == 2 and st["replays"] >= 5, st
    # a batch of replays without host work in between
    g.advance(dt, 20)
    o.advance(dt, 20)
    assert_same("after 30 steps f1", g.get("f1", 2), o.get("f1", 2))
    assert_same("after 30 steps f0", g.get("f0", 2), o.get("f0", 2))
    assert g.graph_stats()["replays"] >= 25
    g.close()
    o.close()


def test_laser_drive_through_device_memory():
    p = two_slab_laser(levels=1)
    say("case: graphs, laser single level fused, 16 steps")
    g, o = va.Run(p, fused=True, graphs=True), OracleRun(p)
    dt = o.calc_dt(0.4)
    for step in range(16):
        g.advance(dt)
        o.advance(dt)
        if step % 4 == 3:
            compare_fields(g, o, f"laser step {step}")
    assert_same("f1", g.get("f1"), o.get("f1"))
    assert np.abs(o.get("Bz")).max() > 0
    assert g.graph_stats()["replays"] >= 10
    g.close()
    o.close()


def test_amr_replays_and_regrid():
    p = two_slab_laser()
    say("case: graphs, three-level laser AMR, 6 steps, regrid, 6 steps")
    g, o = va.Run(p, fused=True, graphs=True), OracleRun(p)
    dt = o.calc_dt(0.4)
    for step in range(6):
        g.advance(dt)
        o.advance(dt)
        compare_mesh(g, o, f"step {step}", what=("f1",))
    assert g.graph_stats()["replays"] >= 4
    g.regrid()
    o.regrid()
    compare_mesh(g, o, "after regrid")
    before = g.graph_stats()["replays"]
    for step in range(6):
        g.advance(dt)
        o.advance(dt)
    compare_mesh(g, o, "after regrid + 6 steps")
    compare_fields(g, o, "after regrid + 6 steps")
    assert g.graph_stats()["replays"] >= before + 4
    g.close()
    o.close()


def test_landau_config_graph_and_eager_agree():
    """Config 1 for 400 steps with graphs and without: the same bits."""
    p = landau()
    say("case: graphs vs eager, Landau 64x256, 400 steps")
    a, b = va.Run(p, graphs=True), va.Run(p, graphs=False)
    dt = a.calc_dt(0.4)
    a.advance(dt, 400)
    b.advance(dt, 400)
    assert_same("f1", a.get("f1"), b.get("f1"))
    assert_same("E", a.get("E"), b.get("E"))
    assert a.graph_stats()["replays"] >= 390 and b.graph_stats()["replays"] == 0
    a.close()
    b.close()
