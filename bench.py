"""Benchmark: phase-space cell-updates/s of the Vlasov-Maxwell time advance.

Workload (one GPU): the relativistic two-stream configuration at its full size,
Nx = Np = 4096 (16.8 M phase-space cells), single unrefined level, FP64, walls
at the domain edges, dt = 0.4 x the CFL bound.  One step = one full 6-stage
Runge-Kutta step of SolverManager::Advance (moments, Poisson, FCT Vlasov stage,
transverse-field stage, six times).  A cell-update is one phase-space cell
advanced by one full step.

Prints one JSON line with the roofline of the dominant kernel (its average
launch time measured live with HIP events on the library stream) and the CPU
baseline (the FP64 CPU oracle on this host's cores, a bounded sample of the
same workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fused|plain]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12           # MI355X HBM3E, B/s
BYTES_PER_UPDATE = 216.0    # algorithmic FP64 bytes per cell-update (stage s: 8(s+1) + 8)


def kernel_bytes_per_cell(kernel, step):
    """Bytes per cell that one launch of `kernel` at RK stage `step` must move
    through its interface, each array touched once."""
    s = step
    if kernel == "k_sub0":
        # read f_s, write FxH_s/FpH_s and FxDS/FpDS, read FxH/FpH of stages < s
        # (at the last stage the partial-sum plane per direction that stage 4
        # writes beside its own fluxes), FxL0/FpL0 written at stage 0 and read
        # afterwards
        return 8 + 16 + 16 + (16 * s if s < 5 else 16) + (16 if s == 4 else 0) + 16
    if kernel == "k_stage":
        # store-flux fused stage: read f_s and f^n, read FxH/FpH of stages < s
        # (at the last stage one partial-sum plane per direction, which stage 4
        # writes in place of its own fluxes), write FxH_s/FpH_s (except the
        # last stage) and f_{s+1}
        return 8 + 8 + (16 * s if s < 5 else 16) + (16 if s < 5 else 0) + 8
    raise ValueError(kernel)


def measured_traffic(kernel, nx, np_):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic_two_stream_{nx}.md")
    if nx != np_ or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


VALU_PEAK = 1024 * 2.4e9 / 4  # wave64 FP64 VALU instructions/s: 256 CUs x 4 SIMDs, 4 cycles each


def measured_valu(kernel, nx, np_, mixed):
    """VALU wave-instructions per launch of `kernel` from the committed
    rocprofv3 SQ_INSTS_VALU summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_valu_two_stream_{nx}.md")
    key = "valu_instructions_per_launch_mixed:" if mixed else "valu_instructions_per_launch:"
    if nx != np_ or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":")[1])
    return None


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """Model, physical cores and threads per core of this host (lscpu)."""
    import subprocess
    info = {"model": None, "physical_cores": os.cpu_count() or 1, "threads_per_core": 1}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=20).stdout
        kv = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        tpc = int(kv.get("Thread(s) per core", "1"))
        cps = int(kv.get("Core(s) per socket", "0"))
        sockets = int(kv.get("Socket(s)", "1"))
        info["threads_per_core"] = tpc
        if cps > 0:
            info["physical_cores"] = cps * sockets
        # the CPUs this process may use (a container's share) bound the cores
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        info["physical_cores"] = max(1, min(info["physical_cores"], avail // max(tpc, 1) or 1))
        info["cpus_available"] = avail
    except Exception:
        pass
    return info


_CPU_CHILD = r"""
import json, sys, time
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
import veritas_amd.params as P
from _oracle import OracleRun
kw = json.loads(sys.argv[2]); workload, steps, reps = kw.pop("workload"), kw.pop("steps"), kw.pop("reps")
params = getattr(P, workload)(**kw)
o = OracleRun(params)
dt = o.calc_dt(0.4)
times = []
for _ in range(reps):
    t0 = time.perf_counter()
    o.advance(dt, steps)
    times.append(time.perf_counter() - t0)
    print(f"bench: cpu baseline run {len(times)}: {times[-1]:.1f} s", file=sys.stderr, flush=True)
print(json.dumps(times))
"""


def cpu_baseline(workload, kw, cells, steps, reps=3, label=""):
    """Time the CPU oracle (test infrastructure, used here only as the
    baseline) on `steps` steps of the same workload, `reps` times in one child
    process after one initialisation; the median is reported.  OpenMP runs one
    thread per physical core, bound close to its core (OMP_PROC_BIND=close,
    OMP_PLACES=cores), set before the child starts."""
    import subprocess
    cpu = host_cpu()
    env = dict(os.environ, OMP_NUM_THREADS=str(cpu["physical_cores"]), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    arg = json.dumps(dict(kw, workload=workload, steps=steps, reps=reps))
    progress(f"cpu baseline: {reps} x {steps} steps on {cpu['physical_cores']} cores")
    r = subprocess.run([sys.executable, "-c", _CPU_CHILD, ROOT, arg], env=env, stdout=subprocess.PIPE,
                       text=True, timeout=900, check=True)
    times = sorted(json.loads(r.stdout.strip().splitlines()[-1]))
    med = times[len(times) // 2]
    return {"value": cells * steps / med, "unit": "cell-updates/s", "cores": cpu["physical_cores"],
            "kind": "port", "cpu_model": cpu["model"], "threads_per_core": cpu["threads_per_core"],
            "omp": "OMP_NUM_THREADS=physical cores, OMP_PROC_BIND=close, OMP_PLACES=cores",
            "runs_s": times,
            "sample": f"{steps} steps of {label}, median of {reps} runs after one initialisation "
                      f"(FP64 CPU restatement, OpenMP; initialisation excluded)"}


def driver_loop_ms(run, dt_max, steps, cfl=0.4):
    """ms per step of the reference driver's loop (veritas.cpp:131-137): dt =
    min(CalculateDt(cfl), dt_max) before every step (a device-to-host read of
    the field state and the bound on the host), then Advance(dt)."""
    run.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.advance(min(run.calc_dt(cfl), dt_max))
    run.sync()
    return (time.perf_counter() - t0) * 1e3 / steps


def mixed_error(va, params, steps):
    """Relative L2 error of the mixed-precision form against the FP64 path
    (bit-identical to the CPU oracle) at 25/50/100 % of `steps`, both taking
    dt = 0.4 x CFL of the FP64 run every step."""
    import numpy as np
    g, r = va.Run(params, mixed=True), va.Run(params)
    n0 = r.get("f1").sum()

    def rel(x, y):
        n = np.linalg.norm(y)
        return float(np.linalg.norm(x - y) / n) if n > 0 else float(np.linalg.norm(x))

    out, step = {}, 0
    for m in (steps // 4, steps // 2, steps):
        while step < m:
            dt = r.calc_dt(0.4)
            g.advance(dt)
            r.advance(dt)
            step += 1
        fg = g.get("f1")
        out[str(m)] = {"f": rel(fg, r.get("f1")), "Ex": rel(g.get("E"), r.get("E")),
                       "Ey": rel(g.get("Ey"), r.get("Ey")), "Bz": rel(g.get("Bz"), r.get("Bz")),
                       "particle_number": abs(fg.sum() / n0 - 1)}
    g.close()
    r.close()
    return {"steps": steps, "reference": "FP64 GPU path (bit-identical to the CPU oracle)",
            "dt": "0.4 x CFL of the FP64 run, every step", "at_step": out}


def amr_mixed_error(va, params, steps):
    """mixed_error for a hierarchy: f over every patch of every level (the
    hierarchy is fixed, so both runs hold the same patches)."""
    import numpy as np
    g, r = va.Run(params, mixed=True), va.Run(params)
    lay = r.layout()

    def rel(x, y):
        n = np.linalg.norm(y)
        return float(np.linalg.norm(x - y) / n) if n > 0 else float(np.linalg.norm(x))

    def all_f(run):
        return np.concatenate([run.patch(int(l), k, "f1").ravel() for l in sorted(set(lay[:, 0]))
                               for k in range(int((lay[:, 0] == l).sum()))])

    out, step = {}, 0
    for m in (steps // 4, steps // 2, steps):
        while step < m:
            dt = r.calc_dt(0.4)
            g.advance(dt)
            r.advance(dt)
            step += 1
        out[str(m)] = {"f": rel(all_f(g), all_f(r)), "Ex": rel(g.get("E"), r.get("E")),
                       "Ey": rel(g.get("Ey"), r.get("Ey")), "Bz": rel(g.get("Bz"), r.get("Bz"))}
    g.close()
    r.close()
    return {"steps": steps, "reference": "FP64 GPU path (bit-identical to the CPU oracle)",
            "dt": "0.4 x CFL of the FP64 run, every step", "at_step": out}


def graph_step_ms(run, dt, steps):
    """ms per step with the steps replayed as recorded graphs (the default
    mode of SolverManager::Advance); the kernel-timed run above launches
    eagerly so that every launch of the dominant kernel carries its events."""
    run.set_graphs(True)
    run.advance(dt, 4)  # each plane state: one eager step, then the recording
    ms = run.time_advance(dt, 0, steps)
    st = run.graph_stats()
    return ms / steps if st["replays"] >= steps else None


def bench_amr(a, va):
    """Laser pulse into underdense plasma: base grid 8192 x 512, three levels at
    ratio 2, p in [-2, 2] m_e c, FP64 (or the mixed-precision form with
    --precision mixed), dt = 0.4 x CFL, sub-step kernels with the device task
    programs.  A cell-update is one cell of one patch of any level
    advanced one full step (nested coarse cells included: they are computed,
    then overwritten by restriction).  The hierarchy is fixed during the timed
    steps (the reference regrids every 20 or more steps)."""
    params = va.laser()
    t0 = time.perf_counter()
    mixed = a.precision == "mixed"
    # --mode plain: every sub-step launched patch by patch (FP64), the plain
    # per-patch form the batched launches are measured against
    batched = a.mode != "plain"
    if not batched and mixed:
        raise SystemExit("bench: the per-patch form (--mode plain) is FP64 only")
    run = va.Run(params, fused=False, graphs=False, mixed=mixed, batched=batched)  # eager: per-launch events
    setup = time.perf_counter() - t0
    progress(f"laser AMR set up in {setup:.1f} s")
    lay = run.layout()
    cells = int((lay[:, 3] * lay[:, 4]).sum())
    dt = run.calc_dt(0.4)
    run.advance(dt, a.warmup)
    run.sync()
    run.timing_begin("k_sub0")
    ms = run.time_advance(dt, 0, a.steps)
    kms, nlaunch = run.timing_end()
    graph_ms = graph_step_ms(run, dt, a.steps)
    progress(f"timed steps: {ms / a.steps:.2f} ms/step")
    # the reference's driver regrids every 22 steps (veritas.cpp:146-151): time
    # three regrids (flags, clustering and, when the boxes change, the rebuild)
    regrid_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        run.regrid()
        run.sync()
        regrid_ms.append((time.perf_counter() - t0) * 1e3)
    regrids, rebuilds = run.regrid_stats()
    value = cells * a.steps / (ms / 1e3)
    per_cell = sum(kernel_bytes_per_cell("k_sub0", s) for s in range(6)) / 6.0
    # k_sub0 launches cover all patches: bytes over all launches / their total time
    achieved = per_cell * cells * 6 * a.steps / (kms / 1e3) if kms else float("nan")
    out = {
        "metric": "phase-space cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64",
        "data": "synthetic (laser pulse into a ramped plasma slab)",
        "config": {"workload": "relativistic laser-plasma, base 8192x512, 3 levels at ratio 2, "
                               + ("mixed precision" if mixed else "FP64") + ", dt=0.4xCFL, fixed hierarchy"
                               + ("" if batched else ", per-patch launches"),
                   "patches_per_level": [int((lay[:, 0] == l).sum()) for l in range(int(lay[:, 0].max()) + 1)],
                   "cells_all_levels": cells, "setup_s": setup, "graph_ms_per_step": graph_ms,
                   "regrid_ms": sorted(regrid_ms)[1], "regrids_that_rebuilt": f"{rebuilds} of {regrids}",
                   "ms_per_step_with_regrid_every_22": ms / a.steps + sorted(regrid_ms)[1] / 22.0},
        "roofline": {"bound": "hbm", "kernel": "k_sub0", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "avg_launch_ms": kms / nlaunch if nlaunch else None, "launches": nlaunch,
                     "kernel_share": kms / ms if ms else None},
    }
    run.close()
    if mixed:
        out["config"]["precision"] = ("FP32 WENO weights and Zalesak ratios in the batched flux and limiter "
                                      "launches; FP64 fluxes, updates, task programs, compensated moments, fields")
        out["error_vs_fp64"] = amr_mixed_error(va, params, a.error_steps) if a.error_steps > 0 else None
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline("laser", {}, cells, a.cpu_steps,
                                           label="the same three-level laser workload")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def serve_pass(va, reqs, arrivals, a):
    srv = va.Server(budget_bytes=int(a.budget_gb * (1 << 30)), max_active=a.slots)
    for k, (r, t) in enumerate(zip(reqs, arrivals)):
        srv.submit(r["params"], r["steps"], arrival=float(t), rid=k)
    t0 = time.perf_counter()
    srv.run()
    wall = time.perf_counter() - t0
    rec, st = srv.records(), srv.stats()
    srv.close()
    return rec, st, wall


def bench_serve(a, va):
    """Config 5: a seeded parameter-scan stream (Landau, two-stream and laser
    runs, Nx 128-4096, Np 256-2048, 1-3 mesh levels; `--req-steps` steps each,
    dt = 0.4 x CFL before every step).  Pass 1 queues every request at t = 0 and
    measures the capacity (solves/s at saturation); pass 2 replays the stream
    with Poisson arrivals at 80 % of that capacity and reports the
    time-to-solution (finish - arrival) percentiles and the solve rate."""
    import numpy as np
    from veritas_amd.serve import poisson_arrivals, scan_requests
    reqs = scan_requests(a.requests, seed=5, steps=a.req_steps)
    print(f"serve: {len(reqs)} requests, pass 1 (saturation)", file=sys.stderr, flush=True)
    rec1, st1, wall1 = serve_pass(va, reqs, np.zeros(len(reqs)), a)
    span1 = max(r["finished"] for r in rec1)
    capacity = len(rec1) / span1
    rate = 0.8 * capacity
    arrivals = poisson_arrivals(len(reqs), rate, seed=7)
    print(f"serve: capacity {capacity:.2f} solves/s, pass 2 at {rate:.2f}/s", file=sys.stderr, flush=True)
    rec, st, wall = serve_pass(va, reqs, arrivals, a)
    lat = np.array([r["finished"] - r["arrival"] for r in rec])
    span = max(r["finished"] for r in rec) - float(arrivals[0])
    cu = sum(r["cell_updates"] for r in rec)
    out = {
        "metric": "scan solves/hr",
        "value": len(rec) / span * 3600.0,
        "unit": "solves/hr",
        "n_gpus": 1,
        "steps": a.req_steps,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded parameter-scan stream)",
        "config": {"workload": f"serving: {len(reqs)} scan requests (Landau / two-stream / laser, Nx 128-4096, "
                               f"Np 256-2048, 1-3 levels, {a.req_steps} steps each), Poisson arrivals at 80% "
                               f"of measured capacity",
                   "budget_gb": a.budget_gb, "slots": a.slots, "capacity_solves_per_s": capacity,
                   "arrival_rate_per_s": rate},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "mean": float(lat.mean()), "max": float(lat.max())},
        "cell_updates_per_s": cu / span,
        "saturation_pass": {"makespan_s": span1, "cell_updates_per_s": sum(r["cell_updates"] for r in rec1) / span1,
                            "stats": st1},
        "stats": st,
        "batched_requests": int(sum(r["batched"] for r in rec)),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["fused", "plain"], default="fused")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 WENO weights and Zalesak ratios, compensated FP64 moments "
                         "(reports its error against the FP64 path at 25/50/100%% of --error-steps)")
    ap.add_argument("--error-steps", type=int, default=500, help="mixed: run length of the error report (0: none)")
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--np", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=2, help="CPU baseline: steps per timed run (3 runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--requests", type=int, default=256, help="serve: requests in the scan stream")
    ap.add_argument("--req-steps", type=int, default=40, help="serve: steps per request")
    ap.add_argument("--budget-gb", type=float, default=32.0, help="serve: device-memory budget")
    ap.add_argument("--slots", type=int, default=16, help="serve: runs stepped at once")
    ap.add_argument("--workload", choices=["two_stream", "laser_amr", "serve"], default="two_stream",
                    help="two_stream: 4096^2 single level (the headline); laser_amr: the "
                         "8192x512 three-level laser configuration")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import veritas_amd as va
    if a.workload == "laser_amr":
        return bench_amr(a, va)
    if a.workload == "serve":
        return bench_serve(a, va)
    params = va.two_stream(nx=a.nx, np=a.np)
    mixed = a.precision == "mixed"
    run = va.Run(params, fused=(a.mode == "fused"), graphs=False, mixed=mixed)  # eager: per-launch events
    dt = run.calc_dt(0.4)
    kernel = "k_stage" if (a.mode == "fused" and va.has_fused_stage()) else "k_sub0"

    run.advance(dt, a.warmup)
    run.sync()
    run.timing_begin(kernel)
    t0 = time.perf_counter()
    ms = run.time_advance(dt, 0, a.steps)
    wall = time.perf_counter() - t0
    kms, nlaunch = run.timing_end()
    graph_ms = graph_step_ms(run, dt, a.steps)

    driver_ms = driver_loop_ms(run, dt, a.steps)  # graphs stay on: the library's default mode

    cells = params.nx * params.np
    value = cells * a.steps / (ms / 1e3)
    # dominant kernel, one launch per RK stage: algorithmic bytes are 216 B per
    # cell-update (stage s reads f_0..f_s and writes f_{s+1}), 36 B per cell and
    # launch on average; the design's own interface bytes (stored fluxes) and
    # the counters' HBM bytes are reported beside them
    algo_bytes_per_launch = BYTES_PER_UPDATE / 6.0 * cells
    design_bytes_per_launch = sum(kernel_bytes_per_cell(kernel, s) for s in range(6)) / 6.0 * cells
    avg_launch_s = (kms / nlaunch) / 1e3 if nlaunch else float("nan")
    achieved = algo_bytes_per_launch / avg_launch_s
    out = {
        "metric": "phase-space cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64",
        "data": "synthetic (seeded two-stream initial condition)",
        "config": {"workload": f"relativistic two-stream, single level, {params.nx}x{params.np} "
                               f"phase-space cells, FP64, dt=0.4xCFL",
                   "nx": params.nx, "np": params.np, "mode": a.mode,
                   "hbm_frac_216B": value * BYTES_PER_UPDATE / HBM_PEAK,
                   "wall_s": wall, "graph_ms_per_step": graph_ms,
                   "ms_per_step_driver_loop": driver_ms,
                   "tile_rows": run.tile_rows() if a.mode == "fused" else None,
                   "tile_waves": run.tile_waves() if a.mode == "fused" else None},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(kernel, params.nx, params.np),
                     "algorithmic_bytes_per_launch": algo_bytes_per_launch,
                     "design_bytes_per_launch": design_bytes_per_launch,
                     "design_frac": design_bytes_per_launch / avg_launch_s / HBM_PEAK,
                     "avg_launch_ms": kms / nlaunch if nlaunch else None,
                     "launches": nlaunch, "kernel_share": kms / ms if ms else None},
    }
    valu = measured_valu(kernel, params.nx, params.np, mixed)
    if valu and avg_launch_s == avg_launch_s:
        # the bound that binds this kernel: FP64 VALU issue (profiles/k_stage_valu_*.md)
        out["valu_roofline"] = {"bound": "valu-issue", "achieved": valu / avg_launch_s / 1e9,
                                "peak": VALU_PEAK / 1e9, "unit": "G wave-instructions/s",
                                "frac": valu / avg_launch_s / VALU_PEAK,
                                "instructions_per_launch": valu}
    run.close()
    if mixed:
        out["config"]["precision"] = "FP32 WENO weights and Zalesak ratios; FP64 fluxes, updates, compensated moments, fields"
        out["error_vs_fp64"] = mixed_error(va, params, a.error_steps) if a.error_steps > 0 else None
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline("two_stream", {"nx": params.nx, "np": params.np}, cells,
                                           a.cpu_steps, label=f"the {params.nx}x{params.np} two-stream workload")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
