#!/usr/bin/env python3
"""Benchmark: FEN fractional-step Navier-Stokes time advance on one MI355X.

Workload: 3-D Taylor-Green vortex, triply periodic 256^3, Re = 1600, FP64,
seeded 1e-3 perturbation, dt from set_timestep (CFL 0.5, U = 2).  One step
is one full navier_stokes_solver call: momentum, divergence, FFT Poisson
solve, projection, pressure update, ghost fills and checks.

Prints one JSON line.  `value` is cell-updates per second (interior cells
advanced one full step per second) with all fields resident in HBM.
`roofline` is the dominant kernel's algorithmic HBM bytes per launch over
its average duration, timed with HIP events recorded on the library's own
stream around that kernel on every tenth step of the timed region; `step_roofline` applies the same
algorithmic model (37 words per cell per step) to the whole step.
`fp32_refine` repeats the measurement in the FP32 form (FP32 fields,
stencils and FFTs; FP64 divergence residual, projection update and
elimination) and reports its speed-up and its relative L2 errors in u, v,
w, p and max|div u| against the FP64 run after the same steps.
`cpu_baseline` times the oracle (this repository's C restatement of the
reference algorithm, OpenMP) on the same workload for a bounded number of
steps on this machine's host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
WORDS_3D = 37           # algorithmic words per cell per step (3-D, ppp)
SAMPLE = 10             # the dominant kernel is timed on every SAMPLE-th step of the timed region


def read_traffic(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated on a copy kernel), as committed in profiles/traffic_<workload>.txt
    (written by fen_amd.rocprof_summary: lines `kernel bytes ...`)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("#"):
            continue
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def copy_probe(nbytes, reps=3):
    """A plain copy kernel of known size (the HIP runtime's device-to-device
    copy kernel): reads nbytes and writes nbytes per launch."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed in copy probe")
    if hip.hipMemset(src, 0, n) != 0:
        raise RuntimeError("hipMemset failed in copy probe")
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, n, 3) != 0:   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed in copy probe")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def cpu_baseline(N, budget_s):
    """Time the oracle on the same workload on this machine's host cores:
    OpenMP over the cores this process may use, then one thread; 1 warm-up
    step and then 5 timed steps each (fewer if a step exceeds the budget),
    median step time."""
    import statistics
    from oracle import oracle as O
    from fen_amd import cases
    from fen_amd._lib import W

    cores = len(os.sched_getaffinity(0))
    case = cases.tgv3d(N, steps=8)

    def make_oracle():
        o = O.Oracle(case.dim, case.n, case.L, bc=case.bc)

        def up(fid, arr):
            if fid <= W:
                o.field(fid)[1:-1, 1:-1, 1:-1] = arr
            else:
                o.field(fid)[...] = arr

        case.apply(o, up)
        return o

    out = {}
    for label, threads in (("all", cores), ("single", 1)):
        O.set_num_threads(threads)
        o = make_oracle()
        t0 = time.perf_counter()
        o.step(1, case.dt)
        first = time.perf_counter() - t0
        nsteps = max(1, min(5, int(budget_s / max(first, 1e-3))))
        times = []
        for s in range(2, 2 + nsteps):
            t0 = time.perf_counter()
            o.step(s, case.dt)
            times.append(time.perf_counter() - t0)
        o.close()
        out[label] = (1e3 * statistics.median(times), nsteps)
    O.set_num_threads(cores)
    ms, n_all = out["all"]
    ms1, n1 = out["single"]
    return {"value": N ** 3 / (ms * 1e-3), "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "ms_per_step": ms,
            "single_thread": {"value": N ** 3 / (ms1 * 1e-3), "ms_per_step": ms1, "cores": 1},
            "sample": f"oracle (C, OpenMP) on tgv3d {N}^3 FP64: 1 warm-up + {n_all} timed steps on {cores} threads "
                      f"and 1 warm-up + {n1} on 1 thread, median step"}


def tgv2d_time(steps=200):
    """2-D 128^2 case: launch-bound, reported as time per step."""
    from fen_amd import cases
    from fen_amd.cases import make_solver as make_gpu
    case = cases.tgv2d(128, steps=steps)
    g = make_gpu(case)
    g.run(0, 10, case.dt)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(10, steps, case.dt)
    g.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    g.close()
    return {"workload": "2-D Taylor-Green vortex 128^2 (pp), FP64", "ms_per_step": ms, "steps": steps}


def case_rate(case, steps=20, precision=0):
    """Time per step of a case (3 warm-up steps, then `steps` steps of fen_ns_run)."""
    from fen_amd.cases import make_solver as make_gpu
    g = make_gpu(case, precision=precision)
    g.run(0, 3, case.dt)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(3, steps, case.dt)
    g.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    g.close()
    cells = case.n[0] * case.n[1] * case.n[2]
    return {"ms_per_step": ms, "value": cells / (ms * 1e-3), "steps": steps}


def channel_rate(steps=20, precision=0):
    """Channel 512x512x256 (x, y periodic, walls at z; ppn: FFT x, y + Thomas z)."""
    from fen_amd import cases
    return case_rate(cases.channel(steps=steps), steps, precision)


def multiphase_rate(steps=100):
    """FP64 time per step of the two-phase mode on the reference's
    Rayleigh-Taylor case (2-D 128 x 512, MF build)."""
    import numpy as np
    from fen_amd import cases
    from fen_amd._lib import VOF
    from fen_amd.cases import make_solver as make_gpu
    case, props, vof = cases.rayleigh_taylor(steps=steps)
    g = make_gpu(case)
    g.set_multiphase(**props)
    full = np.zeros((case.n[0] + 2, case.n[1] + 2, 3), order="F")
    full[1:-1, 1:-1, 1:2] = vof
    g.upload(VOF, full)
    g.update_ghost_nodes(VOF)
    g.update_material_properties()
    dt = g.set_timestep(1.0)
    g.run(0, 3, dt)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(3, steps, dt)
    g.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    md = g.scalars()["maxdiv"]
    g.close()
    cells = case.n[0] * case.n[1]
    return {"workload": "Rayleigh-Taylor 2-D 128x512 (MF, pn)", "ms_per_step": ms, "value": cells / (ms * 1e-3),
            "steps": steps, "maxdiv": md}


def adjoint_rate(N=128, steps=200, every=20):
    """The discrete adjoint on the 3-D periodic 128^3 Re 400 case: 200
    forward steps recorded with checkpoints every 20 (fen_adjoint_record),
    then the backward sweep with recomputation (fen_adjoint_backward, seed
    upload and gradient download included), beside the plain forward run."""
    import numpy as np
    from fen_amd import Solver, cases
    from fen_amd._lib import U, V, W
    from fen_amd.cases import interior_velocity as final_velocity
    from fen_amd.cases import load_state
    case = cases.tgv3d(N, Re=400.0, steps=steps)
    s = Solver(3, case.n, case.L)
    s.set_params(density=1.0, viscosity=case.viscosity, cfl=0.5)
    dt = s.set_timestep(2.0)
    state = dict(case.fields)

    def timed(fn):
        s.synchronize()
        t0 = time.perf_counter()
        out = fn()
        s.synchronize()
        return 1e3 * (time.perf_counter() - t0), out

    load_state(s, state, dt)
    s.run(0, 10, dt)                      # warm-up
    load_state(s, state, dt)
    t_fwd, _ = timed(lambda: s.run(0, steps, dt))
    load_state(s, state, dt)
    t_rec, _ = timed(lambda: s.adjoint_record(0, steps, every, dt))
    seed = [np.asfortranarray(a) for a in final_velocity(s)]
    t_bwd, grads = timed(lambda: s.adjoint_backward(seed))
    s.close()
    assert all(np.isfinite(g).all() for g in grads[:3])
    return {"workload": f"adjoint of tgv3d {N}^3 Re 400, {steps} steps, checkpoints every {every} (FP64)",
            "forward_ms_per_step": t_fwd / steps, "record_ms_per_step": t_rec / steps,
            "backward_ms_per_step": t_bwd / steps, "gradient_over_forward": (t_rec + t_bwd) / t_fwd}


def neumann_rates(steps=20):
    """FP64 time per step with the Neumann-direction solvers at 256^3 (and
    256^2 in 2-D): lid-driven cavity (nnn), duct (npn), 2-D cavity (nn) and
    2-D channel (pn)."""
    from fen_amd import cases
    out = {}
    for name, case in (("cavity3d_256_nnn", cases.cavity3d(256, 256, 256, steps=steps)),
                       ("duct3d_256_npn", cases.duct3d(256, 256, 256, steps=steps)),
                       ("cavity2d_256_nn", cases.cavity2d(256, 256, steps=steps)),
                       ("channel2d_256_pn", cases.channel2d(256, 256, steps=steps))):
        r = case_rate(case, steps, 0)
        r["workload"] = f"{case.name} {'x'.join(str(v) for v in case.n[:case.dim])} ({name.split('_')[-1]})"
        out[name] = r
    return out


def dropin_rate(N=256, steps=100, read_every=1, warmup=5):
    """FEN's navier_stokes_solver as the Fortran layer drives it, one C-ABI
    call sequence per step: fen_set_params (unchanged values), fen_set_dt_o,
    fen_ns_step (asking for maxdiv and maxCFL on the steps the program reads
    them), fen_get_scalars (dt_o, dt_visc, dt_conv; no device sync).
    read_every = 1 is the layer's default (maxdiv, maxCFL current after every
    step); read_every = 100 is `lazy_checks` with print_solver_status every
    100 steps.  3-D Taylor-Green N^3 FP64, wall-clock time per step."""
    import ctypes
    from fen_amd import cases
    from fen_amd.cases import make_solver
    case = cases.tgv3d(N, steps=steps)
    g = make_solver(case)
    L, h, by = g.lib, g.h, ctypes.byref
    gv = (ctypes.c_double * 3)(*case.g)
    dt = ctypes.c_double(case.dt)
    md, mc, dto, dtv, dtc = (ctypes.c_double() for _ in range(5))
    dto.value = case.dt_o

    def step(s):
        rc = L.fen_set_params(h, case.density, case.viscosity, gv, case.cfl, 0)
        rc |= L.fen_set_dt_o(h, dto.value)
        if s % read_every == 0:
            rc |= L.fen_ns_step(h, s, by(dt), by(md), by(mc))
        else:
            rc |= L.fen_ns_step(h, s, by(dt), None, None)
        rc |= L.fen_get_scalars(h, None, None, by(dto), by(dtv), by(dtc))
        if rc:
            raise RuntimeError(f"drop-in step {s}: {L.fen_last_error(h)}")

    for s in range(1, warmup + 1):
        step(s)
    g.synchronize()
    t0 = time.perf_counter()
    for s in range(warmup + 1, warmup + steps + 1):
        step(s)
    g.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    g.close()
    return {"ms_per_step": ms, "value": N ** 3 / (ms * 1e-3), "steps": steps, "checks_read_every": read_every}


def _case(name, steps=20, precision=0):
    from fen_amd import cases
    make = {"cavity3d": lambda: cases.cavity3d(256, 256, 256, steps=steps),
            "duct3d": lambda: cases.duct3d(256, 256, 256, steps=steps),
            "cavity2d": lambda: cases.cavity2d(256, 256, steps=steps),
            "channel2d": lambda: cases.channel2d(256, 256, steps=steps),
            "channel": lambda: cases.channel(steps=steps)}[name]
    return case_rate(make(), steps, precision)


# --case targets: single configurations for profiling runs
CASES = {
    "channel": lambda: _case("channel"),
    "channel_fp32": lambda: _case("channel", precision=1),
    "cavity3d_nnn": lambda: _case("cavity3d"),
    "duct3d_npn": lambda: _case("duct3d"),
    "cavity2d_nn": lambda: _case("cavity2d"),
    "channel2d_pn": lambda: _case("channel2d"),
    "rayleigh_taylor": lambda: multiphase_rate(100),
    "adjoint_128": lambda: adjoint_rate(),
}


def measure(N, steps, warmup, precision):
    """Warm-up, one untimed calibration of the per-kernel breakdown, then the
    timed steps.  Returns (solver, result dict)."""
    from fen_amd import cases
    from fen_amd.cases import make_solver as make_gpu

    case = cases.tgv3d(N, steps=steps)
    g = make_gpu(case, precision=precision)
    dt = case.dt

    # warm-up first: the first launch of each kernel pays code-object loading
    g.run(0, warmup, dt)
    # calibration (untimed): which kernel dominates the step
    g.profile(None)
    g.profile_reset()
    g.run(warmup, 3, dt)
    prof = g.profile_read()
    # the longest launch (each stencil/FFT kernel runs once per step)
    dominant = max(prof, key=lambda k: prof[k]["ms"] / prof[k]["launches"])
    breakdown = {k: round(1e3 * v["ms"] / v["launches"], 2) for k, v in prof.items()}   # us per launch
    g.profile([])

    step0 = 3 + warmup
    # timed region: every SAMPLE-th step, one event just before and one just
    # after the dominant kernel on the library's stream (an event drains the
    # stream, so the other steps run without them)
    if steps > SAMPLE:
        g.profile([dominant])
        g.profile_sample(SAMPLE)
    else:
        # runs of at most SAMPLE steps (profiler passes): every launch of every
        # kernel, timed from the end of the launch before it (the fused run's
        # first step has no projection + momentum launch to sample)
        g.profile(None)
        g.profile_sample(1)
    g.profile_reset()
    g.synchronize()
    t0 = time.perf_counter()
    g.run(step0, steps, dt)
    g.synchronize()
    t1 = time.perf_counter()
    kp = g.profile_read()[dominant]
    g.profile_sample(1)
    g.profile([])
    sc = g.scalars()

    ms = 1e3 * (t1 - t0) / steps
    avg_ms = kp["ms"] / kp["launches"]
    return g, {"ms": ms, "value": N ** 3 / (ms * 1e-3), "dominant": dominant, "avg_ms": avg_ms,
               "bytes": kp["bytes"], "achieved": kp["bytes"] / (avg_ms * 1e-3) / 1e9, "breakdown": breakdown,
               "maxdiv": sc["maxdiv"], "dt": dt, "steps_total": step0 + steps}


def field_errors(g32, g64, dim=3):
    """Relative L2 errors (interior) of the FP32 form against FP64, per
    component and for the velocity field as one vector."""
    import numpy as np
    from fen_amd._lib import P, U, V, W
    out, num, den = {}, 0.0, 0.0
    for name, f in (("u", U), ("v", V), ("w", W), ("p", P)):
        a = g32.download(f)[1:-1, 1:-1, 1:-1]
        b = g64.download(f)[1:-1, 1:-1, 1:-1]
        e2, b2 = float(np.linalg.norm((a - b).ravel())) ** 2, float(np.linalg.norm(b.ravel())) ** 2
        out[name] = (e2 / b2) ** 0.5
        if f != P:
            num, den = num + e2, den + b2
        del a, b
    out["velocity"] = (num / den) ** 0.5
    return out


def fp32_accuracy(N, steps):
    """Both forms over `steps` steps of the Taylor-Green case at N^3 from the
    same seeded state; errors of the FP32 form against FP64."""
    from fen_amd import cases
    from fen_amd._lib import FP32_REFINE
    from fen_amd.cases import make_solver as make_gpu

    case = cases.tgv3d(N, steps=steps)
    res = {}
    for prec in (0, FP32_REFINE):
        g = make_gpu(case, precision=prec)
        g.run(0, steps - 1, case.dt)
        _, md, _ = g.navier_stokes_solver(steps, case.dt)
        res[prec] = (g, md)
    errs = field_errors(res[FP32_REFINE][0], res[0][0])
    return {"workload": f"3-D Taylor-Green vortex {N}^3, Re 1600, {steps} steps", "grid": [N, N, N],
            "steps": steps, "rel_l2_error_vs_fp64": errs, "maxdiv_fp64": res[0][1],
            "maxdiv_fp32_refine": res[FP32_REFINE][1]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="grid size N (N^3)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32+refine form")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in (per-step call sequence) timing")
    ap.add_argument("--no-extra", action="store_true", help="skip the 512^3, channel and 2-D measurements")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU oracle steps")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="BYTES",
                    help="after the run, launch a plain device copy over BYTES bytes three times: "
                         "the known-size reference for rocprofv3 FETCH_SIZE/WRITE_SIZE calibration")
    ap.add_argument("--fp32-accuracy", type=int, nargs=2, metavar=("N", "STEPS"),
                    help="only report the FP32+refine errors against FP64 at N^3 after STEPS steps")
    ap.add_argument("--case", choices=sorted(CASES),
                    help="only time one of the other configurations (the profiling target for rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    if args.fp32_accuracy:
        print(json.dumps(fp32_accuracy(*args.fp32_accuracy)), flush=True)
        return
    if args.case:
        print(json.dumps({args.case: CASES[args.case]()}), flush=True)
        return

    from fen_amd._lib import FP32_REFINE

    N = args.n
    g64, r64 = measure(N, args.steps, args.warmup, 0)
    workload = f"tgv3d_{N}"
    traffic = read_traffic(r64["dominant"], workload)
    step_gbs = WORDS_3D * 8.0 * N ** 3 / (r64["ms"] * 1e-3) / 1e9

    # the drop-in path (the Fortran layer's per-step sequence), next to the run above
    dropin = None
    if not args.no_dropin:
        dropin = {"every_step": dropin_rate(N, args.steps, 1), "every_100_steps": dropin_rate(N, args.steps, 100),
                  "fen_ns_run_ms_per_step": r64["ms"]}
        for v in (dropin["every_step"], dropin["every_100_steps"]):
            v["over_fen_ns_run"] = v["ms_per_step"] / r64["ms"]

    fp32 = None
    if not args.no_fp32:
        g32, r32 = measure(N, args.steps, args.warmup, FP32_REFINE)
        step32 = WORDS_3D * 4.0 * N ** 3 / (r32["ms"] * 1e-3) / 1e9
        fp32 = {"value": r32["value"], "unit": "cell-updates/s", "ms_per_step": r32["ms"], "dtype": "f32+f64",
                "speedup_vs_fp64": r64["ms"] / r32["ms"],
                "roofline": {"bound": "hbm", "kernel": r32["dominant"], "achieved": r32["achieved"],
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r32["achieved"] / HBM_PEAK_GBS,
                             "traffic": read_traffic(r32["dominant"] + ".f32", workload),
                             "bytes_per_launch": r32["bytes"], "avg_launch_ms": r32["avg_ms"]},
                "step_roofline": {"words_per_cell": WORDS_3D, "bytes_per_word": 4, "achieved": step32,
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step32 / HBM_PEAK_GBS},
                "kernel_us_per_launch": r32["breakdown"],
                "error_vs_fp64": dict(steps=r32["steps_total"], rel_l2=field_errors(g32, g64),
                                      maxdiv_fp64=r64["maxdiv"], maxdiv_fp32_refine=r32["maxdiv"])}
        g32.close()

    extra = None
    if not args.no_extra:
        # the other configurations, fewer steps each (same warm-up and timing as above)
        g64.close()
        extra = {}
        g, r512 = measure(512, 20, 3, 0)
        g.close()
        g, r512f = measure(512, 20, 3, FP32_REFINE)
        g.close()
        extra["tgv3d_512"] = {"fp64": {"value": r512["value"], "ms_per_step": r512["ms"]},
                              "fp32_refine": {"value": r512f["value"], "ms_per_step": r512f["ms"]},
                              "speedup_vs_fp64": r512["ms"] / r512f["ms"], "steps": 20}
        extra["channel_512x512x256"] = {"fp64": channel_rate(20, 0), "fp32_refine": channel_rate(20, FP32_REFINE)}
        extra["tgv2d_128"] = tgv2d_time()
        extra["neumann_solvers"] = neumann_rates(20)
        extra["multiphase_2d"] = multiphase_rate(100)
        extra["adjoint_128"] = adjoint_rate()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(N, args.cpu_budget)

    out = {
        "metric": "NS cell-updates/s at 256^3, FP64 vs FP32+refine; % HBM roofline",
        "value": r64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D Taylor-Green vortex {N}^3 triply periodic, Re 1600, FP64",
                   "grid": [N, N, N], "solver": "ppp", "dt": r64["dt"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": r64["dominant"], "achieved": r64["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": r64["achieved"] / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": r64["bytes"], "avg_launch_ms": r64["avg_ms"]},
        "step_roofline": {"words_per_cell": WORDS_3D, "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS},
        "kernel_us_per_launch": r64["breakdown"],
        "kernel_us_note": "kernel_us_per_launch: untimed calibration steps with an event pair around every "
                          "launch (each pair drains the stream, so these read high); roofline.avg_launch_ms "
                          "samples the dominant kernel alone inside the timed region",
        "maxdiv": r64["maxdiv"],
        "dropin": dropin,
        "fp32_refine": fp32,
        "other_configs": extra,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = r64["value"] / cpu["value"]
    if args.no_extra:
        g64.close()
    print(json.dumps(out), flush=True)
    if args.copy_probe > 0:
        copy_probe(int(args.copy_probe))


if __name__ == "__main__":
    main()
