"""veritas_amd: MI355X-native Veritas Vlasov-Maxwell time advance.

Python view of the C++ host library (veritas_amd/lib/libveritas.so), which
keeps the reference's SolverManager/Mesh/Level/Rectangle/EMFieldSolver
interface and runs every numerical step in HIP kernels for gfx950
(veritas_amd/lib/libvx_kernels.so, C ABI in include/veritas_amd.h).

There is no CPU fallback: constructing a `Run` without a usable GPU or without
the built libraries raises.
"""
import ctypes
import os

import numpy as np

from . import serve  # noqa: F401
from .params import (Hooks, ScenarioParams, SettingsDesc, landau, laser,  # noqa: F401
                     two_stream)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
KERNEL_LIB = os.path.join(LIB_DIR, "libvx_kernels.so")
HOST_LIB = os.path.join(LIB_DIR, "libveritas.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "veritas_amd.h")

PLANES = ("f0", "f1", "f2", "FxDS", "FpDS", "Rp", "Rm", "Cx", "Cp", "FxL0", "FpL0")

_host = None


class VeritasError(RuntimeError):
    pass


def build():
    from .build import build as _b
    return _b()


def has_fused_stage():
    """True when the library's walled-patch stage is one fused kernel."""
    L = ctypes.CDLL(KERNEL_LIB)
    return hasattr(L, "vx_fct_stage_fused_available") and L.vx_fct_stage_fused_available() == 1


def host_lib():
    """Load libveritas.so (and through it libvx_kernels.so); raise if absent."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB) or not os.path.exists(KERNEL_LIB):
            raise VeritasError("veritas_amd libraries are not built; run veritas_amd/build.py")
        ctypes.CDLL(KERNEL_LIB, mode=ctypes.RTLD_GLOBAL)
        L = ctypes.CDLL(HOST_LIB)
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.vxh_create.restype = vp
        L.vxh_create.argtypes = [vp, i]
        L.vxh_create_desc.restype = vp
        L.vxh_create_desc.argtypes = [vp, vp, i]
        L.vxh_settings_banner.restype = i
        L.vxh_settings_banner.argtypes = [vp, vp, ctypes.c_char_p, i]
        L.vxh_screen.restype = i
        L.vxh_screen.argtypes = [vp, i, ctypes.c_char_p, i]
        L.vxh_create_error.restype = ctypes.c_char_p
        L.vxh_last_error.restype = ctypes.c_char_p
        L.vxh_last_error.argtypes = [vp]
        L.vxh_destroy.argtypes = [vp]
        L.vxh_calc_dt.argtypes = [vp, d, ctypes.POINTER(d)]
        L.vxh_set_fused.argtypes = [vp, i]
        L.vxh_set_batched.argtypes = [vp, i]
        L.vxh_set_graphs.argtypes = [vp, i]
        L.vxh_set_mixed.argtypes = [vp, i]
        L.vxh_graph_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.vxh_tile_rows.argtypes = [vp, i, i, i]
        L.vxh_tile_waves.argtypes = [vp, i, i, i]
        L.vxh_signatures.restype = i
        L.vxh_signatures.argtypes = [vp, i, i, i, vp, i]
        L.vxh_server_create.restype = vp
        L.vxh_server_create.argtypes = [i, d, i]
        L.vxh_server_destroy.argtypes = [vp]
        L.vxh_server_error.restype = ctypes.c_char_p
        L.vxh_server_error.argtypes = [vp]
        L.vxh_server_submit.argtypes = [vp, i, vp, i, d, d, i]
        L.vxh_server_keep_state.argtypes = [vp, i]
        L.vxh_server_run.argtypes = [vp]
        L.vxh_server_records.restype = i
        L.vxh_server_records.argtypes = [vp, vp, i]
        L.vxh_server_stats.argtypes = [vp, vp]
        L.vxh_server_final_f.restype = i
        L.vxh_server_final_f.argtypes = [vp, i, vp, i]
        L.vxh_advance.argtypes = [vp, d, i]
        L.vxh_advance_fields.argtypes = [vp, d, i]
        L.vxh_sync.argtypes = [vp]
        L.vxh_stage_moments.argtypes = [vp]
        L.vxh_stage_poisson.argtypes = [vp]
        L.vxh_stage_vlasov.argtypes = [vp, d, i]
        L.vxh_stage_vlasov_sub.argtypes = [vp, d, i, i]
        L.vxh_stage_fields.argtypes = [vp, d, i]
        L.vxh_time_advance.argtypes = [vp, d, i, i, ctypes.POINTER(ctypes.c_float)]
        L.vxh_timing_begin.argtypes = [vp, ctypes.c_char_p]
        L.vxh_timing_end.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(i)]
        L.vxh_get.restype = i
        L.vxh_get.argtypes = [vp, ctypes.c_char_p, i, vp, i]
        L.vxh_put.restype = i
        L.vxh_put.argtypes = [vp, ctypes.c_char_p, i, vp, i]
        L.vxh_layout.restype = i
        L.vxh_layout.argtypes = [vp, i, vp, i]
        L.vxh_patch_get.restype = i
        L.vxh_patch_get.argtypes = [vp, i, i, i, ctypes.c_char_p, i, vp, i]
        L.vxh_flags.restype = i
        L.vxh_flags.argtypes = [vp, i, i, vp, i]
        L.vxh_regrid.argtypes = [vp, d]
        L.vxh_program_stats.restype = i
        L.vxh_program_stats.argtypes = [vp, i, vp, i]
        L.vxh_regrid_stats.argtypes = [vp, i, vp]
        L.vxh_file_output.argtypes = [vp, d]
        L.vxh_output_rectangles.argtypes = [vp, d]
        L.vxh_output_flush.argtypes = [vp]
        L.vxh_save.argtypes = [vp, ctypes.c_char_p]
        L.vxh_load.argtypes = [vp, ctypes.c_char_p]
        _host = L
    return _host


def settings_banner(desc, hooks):
    """Settings(desc, hooks) on the host alone (no GPU): its banner
    (Settings.cpp:60-125), or VeritasError with the reason the inputs are
    rejected (Settings.cpp:35-50)."""
    L = host_lib()
    n = L.vxh_settings_banner(ctypes.byref(desc), ctypes.byref(hooks.c), None, 0)
    if n < 0:
        raise VeritasError(L.vxh_create_error().decode())
    buf = ctypes.create_string_buffer(n + 1)
    L.vxh_settings_banner(ctypes.byref(desc), ctypes.byref(hooks.c), buf, n + 1)
    return buf.value.decode()


class _Shape:
    """nx, np (species 0) and cfl of a run made from a SettingsDesc."""

    def __init__(self, desc):
        self.nx, self.np, self.cfl = int(desc.nx), int(desc.np[0]), float(desc.cfl)


class Run:
    """One SolverManager (SolverManager.hpp) on one GPU, for a built-in
    scenario (`Run(params)`) or for the reference's inputs and user hooks
    (`Run.from_desc(desc, hooks)`)."""

    def __init__(self, params, device=0, fused=True, graphs=True, mixed=False, batched=True, _handle=None):
        L = host_lib()
        self.params = params
        self.h = _handle if _handle is not None else L.vxh_create(ctypes.byref(params), device)
        if not self.h:
            raise VeritasError(L.vxh_create_error().decode())
        self.set_fused(fused)
        if not batched:
            self.set_batched(False)
        self.set_graphs(graphs)
        if mixed:
            self.set_mixed(True)

    @classmethod
    def from_desc(cls, desc, hooks, device=0, **kw):
        """A run of the reference's Input/Particles (`SettingsDesc`, SI) and
        user hooks (`Hooks`): Settings(grid, particles, output) followed by
        SolverManager(settings) (veritas.cpp:118-128)."""
        L = host_lib()
        h = L.vxh_create_desc(ctypes.byref(desc), ctypes.byref(hooks.c), device)
        if not h:
            raise VeritasError(L.vxh_create_error().decode())
        run = cls(_Shape(desc), device=device, _handle=h, **kw)
        run._hooks = hooks  # the callbacks must outlive the run
        return run

    def banner(self):
        """The run banner of Settings.cpp:60-125."""
        return self._screen(0)

    def screen(self):
        """The progress table of SolverManager.cpp:9-100 (set-up and regrids)."""
        return self._screen(1)

    def _screen(self, which):
        L = host_lib()
        n = L.vxh_screen(self.h, which, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        L.vxh_screen(self.h, which, buf, n + 1)
        return buf.value.decode()

    def _ck(self, rc):
        if rc != 0:
            raise VeritasError(host_lib().vxh_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            host_lib().vxh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_fused(self, on):
        self._ck(host_lib().vxh_set_fused(self.h, int(bool(on))))

    def set_batched(self, on):
        """False: hierarchies launch every sub-step patch by patch in FP64 (the
        plain per-patch form) instead of one launch per patch table."""
        self._ck(host_lib().vxh_set_batched(self.h, int(bool(on))))

    def set_mixed(self, on):
        """Mixed-precision form (SolverManager::SetMixed): FP32 WENO weights and
        Zalesak ratios in the fused stage of a single walled patch, compensated
        FP64 moment sums; not bit-identical to the FP64 path."""
        self._ck(host_lib().vxh_set_mixed(self.h, int(bool(on))))

    def set_graphs(self, on):
        """Replay steps as recorded hipGraphs (SolverManager::SetGraphs)."""
        self._ck(host_lib().vxh_set_graphs(self.h, int(bool(on))))

    def tile_rows(self, n_x=None, n_p=None):
        """Rows per strip the fused FP64 stage uses at each of the six RK stages
        for an n_x x n_p patch (0 until that stage's first launch times the
        candidates)."""
        n_x = self.params.nx if n_x is None else n_x
        n_p = self.params.np if n_p is None else n_p
        return [host_lib().vxh_tile_rows(self.h, n_x, n_p, s) for s in range(6)]

    def tile_waves(self, n_x=None, n_p=None):
        """Resident waves per SIMD (3 or 4) of the fused stage's kernel form
        chosen with those strips at each RK stage (0 until it is chosen)."""
        n_x = self.params.nx if n_x is None else n_x
        n_p = self.params.np if n_p is None else n_p
        return [host_lib().vxh_tile_waves(self.h, n_x, n_p, s) for s in range(6)]

    def graph_stats(self):
        out = (ctypes.c_int * 2)()
        self._ck(host_lib().vxh_graph_stats(self.h, out))
        return {"replays": out[0], "records": out[1]}

    def calc_dt(self, cfl=None):
        out = ctypes.c_double()
        self._ck(host_lib().vxh_calc_dt(self.h, self.params.cfl if cfl is None else cfl,
                                        ctypes.byref(out)))
        return out.value

    def advance(self, dt, nsteps=1):
        self._ck(host_lib().vxh_advance(self.h, dt, nsteps))

    def advance_fields(self, dt, nsteps=1):
        self._ck(host_lib().vxh_advance_fields(self.h, dt, nsteps))

    def sync(self):
        self._ck(host_lib().vxh_sync(self.h))

    def stage_moments(self):
        self._ck(host_lib().vxh_stage_moments(self.h))

    def stage_poisson(self):
        self._ck(host_lib().vxh_stage_poisson(self.h))

    def stage_vlasov(self, dt, step):
        self._ck(host_lib().vxh_stage_vlasov(self.h, dt, step))

    def stage_vlasov_sub(self, dt, step, sub):
        self._ck(host_lib().vxh_stage_vlasov_sub(self.h, dt, step, sub))

    def stage_fields(self, dt, step):
        self._ck(host_lib().vxh_stage_fields(self.h, dt, step))

    def time_advance(self, dt, warmup, steps):
        ms = ctypes.c_float()
        self._ck(host_lib().vxh_time_advance(self.h, dt, warmup, steps, ctypes.byref(ms)))
        return ms.value

    def timing_begin(self, kernel):
        """Bracket every launch of `kernel` (e.g. "k_sub0") with HIP events on
        the library stream until timing_end()."""
        self._ck(host_lib().vxh_timing_begin(self.h, kernel.encode()))

    def timing_end(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        self._ck(host_lib().vxh_timing_end(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def put(self, what, values, ghosts=0):
        """Write f slot "f0", "f1" or "f2" of the base patch from an array of
        (nx + 2g) x (np + 2g) values; ghost layers beyond g are left alone."""
        a = np.ascontiguousarray(values, dtype=np.float64)
        rc = host_lib().vxh_put(self.h, what.encode(), ghosts, a.ctypes.data, a.size)
        if rc == -2:
            raise KeyError(f"{what} with {a.size} values and {ghosts} ghost layers")
        self._ck(rc)

    def get(self, what, ghosts=0):
        L = host_lib()
        n = L.vxh_get(self.h, what.encode(), ghosts, None, 0)
        if n == -2:
            raise KeyError(what)
        if n < 0:
            self._ck(-1)
        out = np.empty(n, dtype=np.float64)
        self._ck(0 if L.vxh_get(self.h, what.encode(), ghosts, out.ctypes.data, n) == n else -1)
        if what in PLANES or what[:3] in ("FxH", "FpH"):
            p = self.params
            return out.reshape(p.nx + 2 * ghosts, p.np + 2 * ghosts)
        return out

    def scalars(self):
        s = self.get("scalars")
        return dict(Ex0=s[0], time=s[1], dx=s[2], dp=s[3], pmin=s[4], mass=s[5], charge=s[6],
                    n_prepad=int(s[7]))

    # ---- adaptive mesh
    def layout(self, sp=0):
        """Patches as rows (level with 0 the finest, x_pos, p_pos, n_x, n_p,
        up, down, left, right), levels finest first, in the reference's order."""
        L = host_lib()
        n = L.vxh_layout(self.h, sp, None, 0)
        if n < 0:
            self._ck(-1)
        a = (ctypes.c_int * max(n, 1))()
        L.vxh_layout(self.h, sp, a, n)
        return np.array(a[:n], dtype=np.int64).reshape(-1, 9)

    def patch(self, level, idx, what, ghosts=0, sp=0):
        """Plane `what` of patch idx of `level` (0 = finest) as (n_x, n_p)."""
        L = host_lib()
        n = L.vxh_patch_get(self.h, sp, level, idx, what.encode(), ghosts, None, 0)
        if n == -2:
            raise KeyError((level, idx, what))
        if n < 0:
            self._ck(-1)
        out = np.empty(n, dtype=np.float64)
        self._ck(0 if L.vxh_patch_get(self.h, sp, level, idx, what.encode(), ghosts, out.ctypes.data, n) == n
                 else -1)
        lay = self.layout(sp)
        row = lay[lay[:, 0] == level][idx]
        return out.reshape(int(row[3]) + 2 * ghosts, int(row[4]) + 2 * ghosts)

    def signatures(self, level, idx, sp=0):
        """Row and column flag counts of patch `idx` of level `level` (0 =
        coarsest) from the device compaction of the last flags() call."""
        L = host_lib()
        n = L.vxh_signatures(self.h, sp, level, idx, None, 0)
        if n < 0:
            self._ck(-1)
        a = (ctypes.c_int * max(n, 1))()
        L.vxh_signatures(self.h, sp, level, idx, a, n)
        v = np.array(a[:n], dtype=np.int64)
        nr, nc = int(v[2]), int(v[3])
        return {"lo_x": int(v[0]), "lo_p": int(v[1]), "rows": v[4:4 + nr], "cols": v[4 + nr:4 + nr + nc]}

    def flags(self, level, sp=0):
        """Refinement flags of hierarchy level `level` (0 = coarsest) as (i, j) rows."""
        L = host_lib()
        n = L.vxh_flags(self.h, sp, level, None, 0)
        if n < 0:
            self._ck(-1)
        a = (ctypes.c_int * max(n, 1))()
        L.vxh_flags(self.h, sp, level, a, n)
        return np.array(a[:n], dtype=np.int64).reshape(-1, 2)

    def program_stats(self, sp=0):
        """Task batches (one launch each) of species sp's programs:
        {"push": [tasks per batch], "pushc": [...], "flux": [[...] per level]}."""
        L = host_lib()
        n = L.vxh_program_stats(self.h, sp, None, 0)
        if n < 0:
            self._ck(-1)
        a = (ctypes.c_int * max(n, 1))()
        L.vxh_program_stats(self.h, sp, a, n)
        v, k, progs = list(a[:n]), 0, []
        while k < len(v):
            progs.append(v[k + 1:k + 1 + v[k]])
            k += 1 + v[k]
        return {"push": progs[0], "pushc": progs[1], "flux": progs[2:]}

    def regrid_stats(self, sp=0):
        """(regrids since set-up, regrids that rebuilt the patches) of species sp."""
        out = (ctypes.c_int * 2)()
        self._ck(host_lib().vxh_regrid_stats(self.h, sp, out))
        return int(out[0]), int(out[1])

    def regrid(self, t=0.0):
        """SolverManager::reGrid(t) (SolverManager.cpp:48-58)."""
        self._ck(host_lib().vxh_regrid(self.h, float(t)))

    # ---- output files (SolverManager::fileOutput / OutputRectangles)
    def file_output(self, t):
        """Append the field dumps of time t to output/ in the working directory."""
        self._ck(host_lib().vxh_file_output(self.h, t))

    def output_rectangles(self, t):
        """Write output/rectangleData/rectangleData_p*_l*_t<t>.txt."""
        self._ck(host_lib().vxh_output_rectangles(self.h, t))

    def flush_output(self):
        """Wait until every requested file is written."""
        self._ck(host_lib().vxh_output_flush(self.h))

    # ---- checkpoint/restart
    def save(self, path):
        """Write a checkpoint (hierarchy, f, fields, clock) to `path`."""
        self._ck(host_lib().vxh_save(self.h, os.fspath(path).encode()))

    def load(self, path):
        """Replace this run's state with a checkpoint of a run of the same scenario."""
        self._ck(host_lib().vxh_load(self.h, os.fspath(path).encode()))


class Server:
    """Serving loop (host/serve.cpp): requests arrive on the wall clock, are
    admitted by device-memory footprint as slots and memory free up, and are
    stepped together, their single walled patches in shared stage launches.
    `submit(params, steps, arrival)` queues a request (arrival in seconds after
    `run()` starts); `run()` returns when all have finished."""

    def __init__(self, budget_bytes=64 << 30, max_active=16, device=0, keep_state=False):
        L = host_lib()
        self.h = L.vxh_server_create(device, float(budget_bytes), max_active)
        if not self.h:
            raise VeritasError(L.vxh_create_error().decode())
        self._keep = {}
        L.vxh_server_keep_state(self.h, int(bool(keep_state)))
        self._n = 0

    def submit(self, params, steps, arrival=0.0, cfl=0.4, mixed=False, rid=None):
        rid = self._n if rid is None else rid
        self._n += 1
        self._keep[rid] = params  # the struct must outlive the call only; keep the shape for final_f
        host_lib().vxh_server_submit(self.h, rid, ctypes.byref(params), steps, cfl, arrival, int(bool(mixed)))
        return rid

    def run(self):
        L = host_lib()
        if L.vxh_server_run(self.h) != 0:
            raise VeritasError(L.vxh_server_error(self.h).decode())

    def records(self):
        L = host_lib()
        n = L.vxh_server_records(self.h, None, 0)
        a = np.zeros(n)
        L.vxh_server_records(self.h, a.ctypes.data, n)
        keys = ("id", "arrival", "admitted", "started", "finished", "footprint", "cell_updates",
                "batched", "time")
        return [dict(zip(keys, row)) for row in a.reshape(-1, 9)]

    def stats(self):
        out = (ctypes.c_double * 4)()
        host_lib().vxh_server_stats(self.h, out)
        return {"peak_bytes": out[0], "peak_active": int(out[1]), "ticks": int(out[2]),
                "batched_launches": int(out[3])}

    def final_f(self, rid):
        p = self._keep[rid]
        L = host_lib()
        n = L.vxh_server_final_f(self.h, rid, None, 0)
        if n < 0:
            raise KeyError(rid)
        a = np.zeros(n)
        L.vxh_server_final_f(self.h, rid, a.ctypes.data, n)
        return a

    def close(self):
        if getattr(self, "h", None):
            host_lib().vxh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

