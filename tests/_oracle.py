"""ctypes view of the CPU oracle (oracle/_build/libvxoracle.so).

Test infrastructure: the oracle is the checker, never the thing measured."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libvxoracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.vxo_create.restype = vp
        L.vxo_create.argtypes = [vp]
        L.vxo_destroy.argtypes = [vp]
        L.vxo_calc_dt.restype = ctypes.c_double
        L.vxo_calc_dt.argtypes = [vp, ctypes.c_double]
        L.vxo_advance.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.vxo_advance_fields.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.vxo_stage_moments.argtypes = [vp]
        L.vxo_stage_poisson.argtypes = [vp]
        L.vxo_stage_vlasov.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.vxo_stage_vlasov_sub.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.vxo_stage_fields.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.vxo_get.restype = ctypes.c_int
        L.vxo_get.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, vp, ctypes.c_int]
        L.vxo_poisson.argtypes = [vp, vp, vp]
        L.vxo_poisson_approx.argtypes = [vp, vp, vp]
        L.vxo_log.restype = ctypes.c_double
        L.vxo_log.argtypes = [ctypes.c_double]
        d = ctypes.c_double
        L.vxo_weno_nomax.restype = d
        L.vxo_weno_nomax.argtypes = [d, d, d, d, ctypes.c_int]
        L.vxo_face_weno.restype = d
        L.vxo_face_weno.argtypes = [d, d, d, d]
        L.vxo_k1e.restype = ctypes.c_double
        L.vxo_k1e.argtypes = [ctypes.c_double]
        i = ctypes.c_int
        L.vxo_layout.restype = i
        L.vxo_layout.argtypes = [vp, i, vp, i]
        L.vxo_patch_get.restype = i
        L.vxo_patch_get.argtypes = [vp, i, i, i, ctypes.c_char_p, i, vp, i]
        L.vxo_flags.restype = i
        L.vxo_flags.argtypes = [vp, i, i, vp, i]
        L.vxo_regrid.argtypes = [vp]
        L.vxo_set_f.restype = i
        L.vxo_set_f.argtypes = [vp, vp, i]
        L.vxo_vlasov_imposed.argtypes = [vp, ctypes.c_double, i, ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


class OracleRun:
    """One oracle SolverManager for a scenario (species 0, base patch)."""

    def __init__(self, params):
        self.params = params
        self.h = lib().vxo_create(ctypes.byref(params))

    def close(self):
        if self.h:
            lib().vxo_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def calc_dt(self, cfl=None):
        return lib().vxo_calc_dt(self.h, self.params.cfl if cfl is None else cfl)

    def advance(self, dt, nsteps=1):
        lib().vxo_advance(self.h, dt, nsteps)

    def advance_fields(self, dt, nsteps=1):
        """Field-only steps (SolverManager::AdvanceFields): f and J frozen."""
        lib().vxo_advance_fields(self.h, dt, nsteps)

    def get(self, what, ghosts=0):
        n = lib().vxo_get(self.h, what.encode(), ghosts, None, 0)
        if n < 0:
            raise KeyError(what)
        out = np.empty(n, dtype=np.float64)
        lib().vxo_get(self.h, what.encode(), ghosts, out.ctypes.data, n)
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
        lib().vxo_stage_moments(self.h)
        return self.get("charge"), self.get("J")

    def poisson(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        lib().vxo_poisson(self.h, b.ctypes.data, x.ctypes.data)
        return x

    def poisson_approx(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        lib().vxo_poisson_approx(self.h, b.ctypes.data, x.ctypes.data)
        return x

    # ---- adaptive mesh (same accessors as veritas_amd.Run)
    def layout(self, sp=0):
        n = lib().vxo_layout(self.h, sp, None, 0)
        a = (ctypes.c_int * max(n, 1))()
        lib().vxo_layout(self.h, sp, a, n)
        return np.array(a[:n], dtype=np.int64).reshape(-1, 9)

    def patch(self, level, idx, what, ghosts=0, sp=0):
        n = lib().vxo_patch_get(self.h, sp, level, idx, what.encode(), ghosts, None, 0)
        if n < 0:
            raise KeyError((level, idx, what))
        out = np.empty(n, dtype=np.float64)
        lib().vxo_patch_get(self.h, sp, level, idx, what.encode(), ghosts, out.ctypes.data, n)
        lay = self.layout(sp)
        row = lay[lay[:, 0] == level][idx]
        return out.reshape(int(row[3]) + 2 * ghosts, int(row[4]) + 2 * ghosts)

    def flags(self, level, sp=0):
        n = lib().vxo_flags(self.h, sp, level, None, 0)
        a = (ctypes.c_int * max(n, 1))()
        lib().vxo_flags(self.h, sp, level, a, n)
        return np.array(a[:n], dtype=np.int64).reshape(-1, 2)

    def regrid(self):
        lib().vxo_regrid(self.h)

    # ---- imposed-field runs (pins of the Vlasov advance by exact solutions)
    def set_f(self, f):
        """Set slots 0 and 1 of the base patch's interior to `f` (nx x np)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        if lib().vxo_set_f(self.h, f.ctypes.data, f.size) != 0:
            raise ValueError("f has the wrong size")

    def vlasov_imposed(self, dt, nsteps, E0=0.0, omega=0.0):
        """Vlasov steps with the uniform field E0 cos(omega t) imposed at every
        stage (t = the stage clock)."""
        lib().vxo_vlasov_imposed(self.h, dt, nsteps, E0, omega)
