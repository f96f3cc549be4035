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
stream inside the timed region; `step_roofline` applies the same
algorithmic model (37 words per cell per step) to the whole step.
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
    """Time the oracle on the same workload: 1 warm-up step, then as many
    steps as fit the budget (at least 1, at most 5)."""
    from oracle import oracle as O
    from fen_amd import cases
    from tests.common import make_oracle

    case = cases.tgv3d(N, steps=8)
    o = make_oracle(case)
    t0 = time.perf_counter()
    o.step(1, case.dt)
    first = time.perf_counter() - t0
    nsteps = max(1, min(5, int(budget_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    for s in range(2, 2 + nsteps):
        o.step(s, case.dt)
    el = time.perf_counter() - t0
    ms = 1e3 * el / nsteps
    cores = O.num_threads()
    o.close()
    return {"value": N ** 3 / (ms * 1e-3), "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "ms_per_step": ms,
            "sample": f"oracle (C, OpenMP {cores} threads) on tgv3d {N}^3 FP64: 1 warm-up + {nsteps} timed steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="grid size N (N^3)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU oracle steps")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="BYTES",
                    help="after the run, launch a plain device copy over BYTES bytes three times: "
                         "the known-size reference for rocprofv3 FETCH_SIZE/WRITE_SIZE calibration")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    from fen_amd import cases
    from tests.common import make_gpu

    N = args.n
    case = cases.tgv3d(N, steps=args.steps)
    g = make_gpu(case)
    dt = case.dt

    # warm-up first: the first launch of each kernel pays code-object loading
    g.run(0, args.warmup, dt)
    # calibration (untimed): which kernel dominates the step
    g.profile(None)
    g.profile_reset()
    g.run(args.warmup, 3, dt)
    prof = g.profile_read()
    # the longest launch (each stencil/FFT kernel runs once per step)
    dominant = max(prof, key=lambda k: prof[k]["ms"] / prof[k]["launches"])
    breakdown = {k: round(1e3 * v["ms"] / v["launches"], 2) for k, v in prof.items()}   # us per launch
    pred = g.profile_predecessor(dominant)
    g.profile([])

    step0 = 3 + args.warmup
    # timed region: one event after the dominant kernel and one after the
    # launch preceding it, so its duration is measured on the library's stream
    g.profile([dominant] + ([pred] if pred and pred != dominant else []))
    g.profile_reset()
    g.synchronize()
    t0 = time.perf_counter()
    g.run(step0, args.steps, dt)
    g.synchronize()
    t1 = time.perf_counter()
    kp = g.profile_read()[dominant]
    g.profile([])
    sc = g.scalars()

    ms = 1e3 * (t1 - t0) / args.steps
    value = N ** 3 / (ms * 1e-3)
    avg_ms = kp["ms"] / kp["launches"]
    achieved = kp["bytes"] / (avg_ms * 1e-3) / 1e9
    workload = f"tgv3d_{N}_fp64"
    traffic = read_traffic(dominant, workload)
    step_gbs = WORDS_3D * 8.0 * N ** 3 / (ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(N, args.cpu_budget)

    out = {
        "metric": "NS cell-updates/s at 256^3, FP64 vs FP32+refine; % HBM roofline",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D Taylor-Green vortex {N}^3 triply periodic, Re 1600, FP64",
                   "grid": [N, N, N], "solver": "ppp", "dt": dt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": kp["bytes"], "avg_launch_ms": avg_ms},
        "step_roofline": {"words_per_cell": WORDS_3D, "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS},
        "kernel_us_per_launch": breakdown,
        "maxdiv": sc["maxdiv"],
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    if args.copy_probe > 0:
        copy_probe(int(args.copy_probe))


if __name__ == "__main__":
    main()
