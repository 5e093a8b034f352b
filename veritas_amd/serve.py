"""Parameter-scan request streams for the serving loop (veritas_amd.Server).

`scan_requests(n, seed)` draws n requests: a Landau run (k in [2, 6], theta in
[0.005, 0.02]), a two-stream run (drift in [0.1, 0.3]) or a laser run (a0 in
[0.2, 1.0], plateau density in [0.005, 0.02] n_c, 1 to 3 mesh levels), on
Nx in {128, ..., 4096} and Np in {256, ..., 2048}.  `poisson_arrivals` gives
arrival times of a Poisson process of a given rate."""
import math

import numpy as np

from . import params as P

NX = (128, 256, 512, 1024, 2048, 4096)
NP = (256, 512, 1024, 2048)


def scan_requests(n=256, seed=5, steps=40):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        kind = int(rng.integers(3))
        nx = int(rng.choice(NX))
        np_ = int(rng.choice(NP))
        if kind == 0:
            kw = float(rng.uniform(2.0, 6.0))
            p = P.landau(nx=nx, np=np_, k=kw, length=2 * math.pi / kw, theta=float(rng.uniform(0.005, 0.02)))
            desc = f"landau k={kw:.2f}"
        elif kind == 1:
            drift = float(rng.uniform(0.1, 0.3))
            p = P.two_stream(nx=nx, np=np_, drift=drift, pmin=-2 * drift - 0.2, pmax=2 * drift + 0.2)
            desc = f"two-stream drift={drift:.2f}"
        else:
            levels = int(rng.integers(1, 4))
            waves = nx / 16.0  # 16 base cells per laser wavelength
            p = P.laser(nx=nx, np=np_, levels=levels, wavelengths=waves, ramp=waves / 4,
                        a0=float(rng.uniform(0.2, 1.0)), n_over_nc=float(rng.uniform(0.005, 0.02)))
            desc = f"laser levels={levels}"
        out.append({"params": p, "steps": steps, "kind": desc, "cells": nx * np_})
    return out


def poisson_arrivals(n, rate, seed=7):
    """Arrival times (s) of a Poisson process with `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
