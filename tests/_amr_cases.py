"""Adaptive-mesh test scenarios shared by the oracle and GPU parity tests.

A laser pulse entering a plasma split into two slabs by a vacuum gap: the
refinement flags fall in two separate x-ranges and the high minimum
efficiency makes the clustering cut them into many small patches, so the
ghost fill, limiter-C sync and coarse-fine flux paths between same-level and
coarse-fine neighbours are all exercised."""
import numpy as np

import veritas_amd.params as P


def two_slab_laser(**kw):
    d = dict(nx=128, np=64, wavelengths=40.0, ramp=5.0, gap=8.0, pmin=-0.5, pmax=0.5, theta=0.01,
             refinement_criteria=0.01, min_efficiency=0.85, levels=3, size_weight=0.0)
    d.update(kw)
    return P.laser(**d)


def particle_number(run, dx0, dp0, r=2):
    """Sum of f1 * dx * dp over the cells of every level not covered by a
    finer patch (levels finest first; level l has cells r^-depth smaller)."""
    lay = run.layout()
    total = 0.0
    nlev = int(lay[:, 0].max()) + 1 if len(lay) else 0
    maxd = nlev - 1
    for lvl in range(nlev):
        rows = lay[lay[:, 0] == lvl]
        scale = float(r) ** (lvl - maxd)  # depth lvl: cells r^(lvl - maxDepth) of the coarsest
        for k in range(len(rows)):
            f = run.patch(lvl, k, "f1")
            nested = run.patch(lvl, k, "nested") if lvl > 0 else np.zeros_like(f)
            total += float(np.sum(f[nested == 0])) * (dx0 * scale) * (dp0 * scale)
    return total
