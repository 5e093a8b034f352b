"""vx_poisson_noalloc (include/veritas_amd.h) against vx_poisson, through the
C ABI: the same inputs give the same phi, E and Ex0 bit for bit, on two calls
in a row (the second with a smaller n_f, so the fixed work array holds the
first call's values when the second starts), with no workspace passed
(fs->work = NULL); sizes beyond the array's limit are rejected."""
import ctypes

import numpy as np
import pytest

import veritas_amd as va

pytestmark = pytest.mark.gpu

D = ctypes.c_double
P = ctypes.c_void_p


class FieldState(ctypes.Structure):  # vx_field_state
    _fields_ = [("n_f", ctypes.c_int32), ("n_prepad", ctypes.c_int32), ("n_postpad", ctypes.c_int32),
                ("n_e", ctypes.c_int32), ("dx_f", D), ("dx2", D), ("field_coef", D)] + \
               [(n, P) for n in ("charge", "J", "neutral", "phi", "E", "Ex0", "By", "Bz", "Ey", "Ez",
                                 "Ay", "Az", "a2", "work")]


def kernels():
    va.host_lib()  # loads libvx_kernels.so too
    L = ctypes.CDLL(va.KERNEL_LIB)
    L.vx_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
    L.vx_ctx_destroy.argtypes = [P]
    L.vx_last_error.restype = ctypes.c_char_p
    L.vx_last_error.argtypes = [P]
    L.vx_malloc.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(P)]
    L.vx_free.argtypes = [P, P]
    L.vx_memcpy_h2d.argtypes = [P, P, P, ctypes.c_size_t]
    L.vx_memcpy_d2h.argtypes = [P, P, P, ctypes.c_size_t]
    L.vx_sync.argtypes = [P]
    L.vx_field_work_elems.restype = ctypes.c_size_t
    L.vx_field_work_elems.argtypes = [ctypes.c_int32]
    for f in (L.vx_poisson, L.vx_poisson_noalloc):
        f.argtypes = [P, ctypes.POINTER(FieldState)]
    return L


class State:
    """A field state of n_f cells in device memory, with the given charge,
    neutralising charge and Ex0; `work` only when asked for."""

    def __init__(self, L, ctx, n_f, charge, neutral, ex0, work):
        self.L, self.ctx, self.bufs = L, ctx, []
        ne = n_f + 4
        dx = 1.0e-7
        self.fs = FieldState(n_f=n_f, n_prepad=2, n_postpad=2, n_e=ne, dx_f=dx, dx2=dx ** 2.0,
                             field_coef=1.0 / (12 * dx))
        sizes = dict(charge=n_f, J=n_f, neutral=n_f, phi=n_f, E=n_f, Ex0=1, By=8 * ne, Bz=8 * ne, Ey=8 * ne,
                     Ez=8 * ne, Ay=8 * ne, Az=8 * ne, a2=n_f + 1)
        if work:
            sizes["work"] = L.vx_field_work_elems(n_f)
        for name, n in sizes.items():
            setattr(self.fs, name, self.alloc(n))
        self.put("charge", charge)
        self.put("neutral", neutral)
        self.put("Ex0", np.array([ex0]))

    def alloc(self, n):
        p = P()
        assert self.L.vx_malloc(self.ctx, 8 * n, ctypes.byref(p)) == 0
        self.bufs.append(p)
        return p

    def put(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.L.vx_memcpy_h2d(self.ctx, getattr(self.fs, name), a.ctypes.data, a.nbytes) == 0

    def get(self, name, n):
        a = np.empty(n)
        assert self.L.vx_memcpy_d2h(self.ctx, a.ctypes.data, getattr(self.fs, name), a.nbytes) == 0
        return a

    def free(self):
        for p in self.bufs:
            self.L.vx_free(self.ctx, p)


def test_noalloc_poisson_matches_the_workspace_entry_point():
    L = kernels()
    ctx = P()
    assert L.vx_ctx_create(0, ctypes.byref(ctx)) == 0
    rng = np.random.default_rng(17)
    try:
        for n_f in (20000, 6000, 2048):  # a larger call, then smaller ones (2048: one workgroup)
            print(f"case: vx_poisson_noalloc vs vx_poisson, n_f = {n_f}", flush=True)
            charge = rng.normal(0.0, 1.0e5, n_f)
            neutral = -np.roll(charge, 3) * 0.999
            ex0 = float(rng.normal(0.0, 1e3))
            a = State(L, ctx, n_f, charge, neutral, ex0, work=True)
            b = State(L, ctx, n_f, charge, neutral, ex0, work=False)
            assert b.fs.work is None
            assert L.vx_poisson(ctx, ctypes.byref(a.fs)) == 0, L.vx_last_error(ctx)
            rc = L.vx_poisson_noalloc(ctx, ctypes.byref(b.fs))
            assert rc == 0, L.vx_last_error(ctx)
            assert L.vx_sync(ctx) == 0
            for name, n in (("phi", n_f), ("E", n_f), ("Ex0", 1)):
                ga, gb = a.get(name, n), b.get(name, n)
                assert np.array_equal(ga, gb), (n_f, name, np.abs(ga - gb).max())
            assert np.abs(a.get("phi", n_f)).max() > 0
            a.free()
            b.free()
        big = State(L, ctx, 65538, np.zeros(65538), np.zeros(65538), 0.0, work=False)
        assert L.vx_poisson_noalloc(ctx, ctypes.byref(big.fs)) != 0
        assert b"exceeds its fixed work array" in L.vx_last_error(ctx)
        big.free()
    finally:
        L.vx_ctx_destroy(ctx)
