"""The config-5 request stream (veritas_amd.serve) is drawn from its seed and
stays inside the stated parameter ranges."""
import numpy as np

from veritas_amd.serve import NP, NX, poisson_arrivals, scan_requests


def test_stream_is_seeded_and_in_range():
    a, b = scan_requests(64, seed=3), scan_requests(64, seed=3)
    assert [x["kind"] for x in a] == [x["kind"] for x in b]
    for r in a:
        p = r["params"]
        assert p.nx in NX and p.np in NP
        if r["kind"].startswith("landau"):
            assert 2.0 <= p.k <= 6.0 and 0.005 <= p.theta <= 0.02
        elif r["kind"].startswith("two-stream"):
            assert 0.1 <= p.drift <= 0.3
        else:
            assert 0.2 <= p.a0 <= 1.0 and 0.005 <= p.n_over_nc <= 0.02 and 1 <= p.levels <= 3
    assert len({r["kind"].split()[0] for r in a}) == 3


def test_poisson_arrivals_have_the_rate():
    t = poisson_arrivals(20000, rate=50.0, seed=1)
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 50.0) < 2.0
