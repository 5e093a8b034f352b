"""Host logic of the adaptive mesh, on the CPU.

The nesting list of Mesh::updateHierarchy (Mesh.cpp:315-339,
mergeDownFlaggedData): the level cells under a finer box, which the host forms
as the rectangle of truncated index ranges instead of visiting every fine cell.
Checked against the reference's loop, restated below, on random boxes, level
sizes and refinement ratios (including boxes that reach past the level's edges)."""
import ctypes

import numpy as np

import veritas_amd as va


def reference_cells_under(nmax, pmax, fx, fp, i0, i1, j0, j1):
    out = set()
    for i in range(i0, i1 + 1):
        for j in range(j0, j1 + 1):
            ic, jc = int(i * fx), int(j * fp)  # C++ truncation toward zero
            if -1 < ic < nmax and -1 < jc < pmax:
                out.add((ic, jc))
    return sorted(out)


def host_cells_under(nmax, pmax, fx, fp, i0, i1, j0, j1):
    L = va.host_lib()
    f = L.vxh_cells_under_box
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 4 + \
        [ctypes.c_void_p, ctypes.c_int]
    n = f(nmax, pmax, fx, fp, i0, i1, j0, j1, None, 0)
    out = np.zeros(2 * max(n, 1), dtype=np.int32)
    assert f(nmax, pmax, fx, fp, i0, i1, j0, j1, out.ctypes.data, n) == n
    return [tuple(int(x) for x in out[2 * k:2 * k + 2]) for k in range(n)]


def test_nesting_list_equals_the_reference_loop():
    rng = np.random.default_rng(11)
    for _ in range(300):
        r = int(rng.integers(2, 5))
        nmax, pmax = int(rng.integers(8, 120)), int(rng.integers(8, 90))
        nf, pf = nmax * r * r, pmax * r * r  # the level two finer
        fx, fp = nmax / float(nf), pmax / float(pf)
        a0 = int(rng.integers(-3 * r, nf)) - 2 * r
        b0 = int(rng.integers(-3 * r, pf)) - 2 * r
        a1 = a0 + int(rng.integers(0, 60))
        b1 = b0 + int(rng.integers(0, 60))
        got = host_cells_under(nmax, pmax, fx, fp, a0, a1, b0, b1)
        want = reference_cells_under(nmax, pmax, fx, fp, a0, a1, b0, b1)
        assert got == want, (nmax, pmax, r, a0, a1, b0, b1)
