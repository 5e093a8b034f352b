"""ctypes view of the CPU oracle (oracle/_build/libvxoracle.so).

Test infrastructure: the oracle is the checker, never the thing measured."""
import contextlib
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libvxoracle.so")

_lib = None


def bind(L):
    """Declare the C signatures of an oracle library loaded with ctypes."""
    vp = ctypes.c_void_p
    d = ctypes.c_double
    i = ctypes.c_int
    L.vxo_create.restype = vp
    L.vxo_create.argtypes = [vp]
    L.vxo_create_desc.restype = vp
    L.vxo_create_desc.argtypes = [vp, vp]
    L.vxo_destroy.argtypes = [vp]
    L.vxo_calc_dt.restype = d
    L.vxo_calc_dt.argtypes = [vp, d]
    L.vxo_advance.argtypes = [vp, d, i]
    L.vxo_advance_fields.argtypes = [vp, d, i]
    L.vxo_stage_moments.argtypes = [vp]
    L.vxo_stage_poisson.argtypes = [vp]
    L.vxo_stage_vlasov.argtypes = [vp, d, i]
    L.vxo_stage_vlasov_sub.argtypes = [vp, d, i, i]
    L.vxo_stage_fields.argtypes = [vp, d, i]
    L.vxo_get.restype = i
    L.vxo_get.argtypes = [vp, ctypes.c_char_p, i, vp, i]
    L.vxo_poisson.argtypes = [vp, vp, vp]
    L.vxo_poisson_approx.argtypes = [vp, vp, vp]
    L.vxo_log.restype = d
    L.vxo_log.argtypes = [d]
    L.vxo_weno_nomax.restype = d
    L.vxo_weno_nomax.argtypes = [d, d, d, d, i]
    L.vxo_face_weno.restype = d
    L.vxo_face_weno.argtypes = [d, d, d, d]
    L.vxo_k1e.restype = d
    L.vxo_k1e.argtypes = [d]
    L.vxo_layout.restype = i
    L.vxo_layout.argtypes = [vp, i, vp, i]
    L.vxo_patch_get.restype = i
    L.vxo_patch_get.argtypes = [vp, i, i, i, ctypes.c_char_p, i, vp, i]
    L.vxo_flags.restype = i
    L.vxo_flags.argtypes = [vp, i, i, vp, i]
    L.vxo_regrid.argtypes = [vp]
    L.vxo_set_f.restype = i
    L.vxo_set_f.argtypes = [vp, vp, i]
    L.vxo_vlasov_imposed.argtypes = [vp, d, i, d, d]
    L.vxo_interp.restype = i
    L.vxo_interp.argtypes = [vp, i, vp, vp, i]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = bind(ctypes.CDLL(LIB_PATH))
    return _lib


@contextlib.contextmanager
def library(path):
    """Run the oracle from another build of it (a copy with a planted
    mistake, tests/test_oracle_plants.py) inside the `with` block."""
    global _lib
    saved = lib()
    _lib = bind(ctypes.CDLL(path))
    try:
        yield _lib
    finally:
        _lib = saved


class _Shape:
    def __init__(self, desc):
        self.nx, self.np, self.cfl = int(desc.nx), int(desc.np[0]), float(desc.cfl)


class OracleRun:
    """One oracle SolverManager for a built-in scenario, or (from_desc) for the
    reference's inputs and user hooks; get() reads species 0's base patch."""

    def __init__(self, params, _handle=None):
        self.params = params
        self.L = lib()  # the library this run lives in
        self.h = _handle if _handle is not None else self.L.vxo_create(ctypes.byref(params))

    @classmethod
    def from_desc(cls, desc, hooks):
        h = lib().vxo_create_desc(ctypes.byref(desc), ctypes.byref(hooks.c))
        if not h:
            raise ValueError("the oracle rejected the inputs")
        run = cls(_Shape(desc), _handle=h)
        run._hooks = hooks
        return run

    def close(self):
        if self.h:
            self.L.vxo_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def calc_dt(self, cfl=None):
        return self.L.vxo_calc_dt(self.h, self.params.cfl if cfl is None else cfl)

    def advance(self, dt, nsteps=1):
        self.L.vxo_advance(self.h, dt, nsteps)

    def advance_fields(self, dt, nsteps=1):
        """Field-only steps (SolverManager::AdvanceFields): f and J frozen."""
        self.L.vxo_advance_fields(self.h, dt, nsteps)

    def get(self, what, ghosts=0):
        n = self.L.vxo_get(self.h, what.encode(), ghosts, None, 0)
        if n < 0:
            raise KeyError(what)
        out = np.empty(n, dtype=np.float64)
        self.L.vxo_get(self.h, what.encode(), ghosts, out.ctypes.data, n)
        if what in ("f0", "f1", "f2", "FxDS", "FpDS", "Rp", "Rm", "Cx", "Cp") or what[:3] in (
                "FxH", "FpH", "FxL", "FpL"):
            p = self.params
            return out.reshape(p.nx + 2 * ghosts, p.np + 2 * ghosts)
        return out

    def scalars(self):
        s = self.get("scalars")
        return dict(Ex0=s[0], time=s[1], dx=s[2], dp=s[3], pmin=s[4], mass=s[5], charge=s[6],
                    n_prepad=int(s[7]))

    def moments(self):
        self.L.vxo_stage_moments(self.h)
        return self.get("charge"), self.get("J")

    def poisson(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        self.L.vxo_poisson(self.h, b.ctypes.data, x.ctypes.data)
        return x

    def poisson_approx(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        self.L.vxo_poisson_approx(self.h, b.ctypes.data, x.ctypes.data)
        return x

    # ---- adaptive mesh (same accessors as veritas_amd.Run)
    def layout(self, sp=0):
        n = self.L.vxo_layout(self.h, sp, None, 0)
        a = (ctypes.c_int * max(n, 1))()
        self.L.vxo_layout(self.h, sp, a, n)
        return np.array(a[:n], dtype=np.int64).reshape(-1, 9)

    def patch(self, level, idx, what, ghosts=0, sp=0):
        n = self.L.vxo_patch_get(self.h, sp, level, idx, what.encode(), ghosts, None, 0)
        if n < 0:
            raise KeyError((level, idx, what))
        out = np.empty(n, dtype=np.float64)
        self.L.vxo_patch_get(self.h, sp, level, idx, what.encode(), ghosts, out.ctypes.data, n)
        lay = self.layout(sp)
        row = lay[lay[:, 0] == level][idx]
        return out.reshape(int(row[3]) + 2 * ghosts, int(row[4]) + 2 * ghosts)

    def flags(self, level, sp=0):
        n = self.L.vxo_flags(self.h, sp, level, None, 0)
        a = (ctypes.c_int * max(n, 1))()
        self.L.vxo_flags(self.h, sp, level, a, n)
        return np.array(a[:n], dtype=np.int64).reshape(-1, 2)

    def regrid(self):
        self.L.vxo_regrid(self.h)

    def interp(self, which, f5):
        """GetInterpolantsREF (which 0) or GetInterpolantsREL (1) of the base
        patch from the cell averages of cells -2..2."""
        f = np.ascontiguousarray(f5, dtype=np.float64)
        n = self.L.vxo_interp(self.h, which, f.ctypes.data, None, 0)
        out = np.empty(n)
        self.L.vxo_interp(self.h, which, f.ctypes.data, out.ctypes.data, n)
        return out

    # ---- imposed-field runs (pins of the Vlasov advance by exact solutions)
    def set_f(self, f):
        """Set slots 0 and 1 of the base patch's interior to `f` (nx x np)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        if self.L.vxo_set_f(self.h, f.ctypes.data, f.size) != 0:
            raise ValueError("f has the wrong size")

    def vlasov_imposed(self, dt, nsteps, E0=0.0, omega=0.0):
        """Vlasov steps with the uniform field E0 cos(omega t) imposed at every
        stage (t = the stage clock)."""
        self.L.vxo_vlasov_imposed(self.h, dt, nsteps, E0, omega)
