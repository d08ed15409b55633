#!/usr/bin/env python3
"""Throughput of the RK3 split-explicit dynamics step on one MI355X.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5]

Workload: the 64x64x40 dry warm bubble (1 km grid, 20 km top, periodic, damp_opt 3 Rayleigh layer of
5 km, dt 6 s with 6 acoustic steps, 5th/3rd-order advection), plain FP64, synthetic seeded initial state
built on the host.  A step is one RK3 large step (three stages, ten acoustic substeps); the timed region
is K steps with the state already resident in HBM, bracketed by device synchronisations.

Prints one JSON line on stdout with the throughput in grid-point-steps/s (nx*ny*nz*steps/s), the
roofline of the dominant kernel and the CPU baseline (the C++/OpenMP oracle on this host's cores, a
bounded sample of the same workload).

The timed region replays the step's HIP graph K times.  The dominant kernel's launch duration comes from
HIP events recorded on the domain's stream around each of its launches during K further steps run
launch by launch (events cannot be read back from inside a replayed graph); its algorithmic bytes are
the distinct arrays it reads and writes x 8 B x the points of one launch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "tests"))

METRIC = "grid-point-steps/s FP32-batched vs FP64 plain at stated L2; p50/p99 forecast latency"
HBM_PEAK_GBS = 8000.0
# kernels whose event timing is sampled to find the dominant one
CANDIDATES = ["large_prep", "tend_u", "tend_v", "tend_w", "tend_t", "tend_ph", "ss_prep", "ss_column", "uv", "acoustic",
              "finish"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(nx=64, ny=64, nz=40):
    from cppwrf_amd.dyn import Case
    case = Case.bubble3d(nx, ny, nz, dx=1000.0, ztop=20000.0, dt=6.0, seed=1, noise_amp=0.05)
    prm = case.params(time_step_sound=6, damp_opt=3, zdamp=5000.0, dampcoef=0.2)
    return case, prm


def profile(rk, name, steps):
    import ctypes as C
    from cppwrf_amd._abi import check, lib
    check(lib().mi_profile_kernel(rk.handle, name.encode()))
    rk.step(steps)
    ms, n, b = C.c_double(), C.c_int(), C.c_double()
    check(lib().mi_profile_read(rk.handle, C.byref(ms), C.byref(n), C.byref(b)))
    check(lib().mi_profile_kernel(rk.handle, None))
    return ms.value, n.value, b.value


def cpu_baseline(case, prm, budget_s=12.0):
    """The oracle (C++/OpenMP, bit-identical to the reference CPU FP64 build) on this host's cores."""
    from oracle_lib import OracleDomain, olib
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    olib().orc_set_threads(threads)
    orc = OracleDomain(case, prm)
    t = time.perf_counter()
    orc.step(1)
    one = time.perf_counter() - t
    n = max(1, min(100, int(budget_s / max(one, 1e-3))))
    t = time.perf_counter()
    orc.step(n)
    dt = time.perf_counter() - t
    g = case.grid
    return {"value": g.nx * g.ny * g.nz * n / dt, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle, {n} steps of the same 64x64x40 workload after 1 untimed step, {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default block shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (each forecast domain runs on a "
              f"single GPU; independent domains shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)

    from cppwrf_amd.dyn import Context, RK3

    case, prm = workload()
    g = case.grid
    ctx = Context(0)
    rk = RK3(ctx, g, prm, case)
    if not args.no_autotune:
        log("autotuned block shapes:", rk.autotune(1))  # steps the domain, then restores its state
    rk.step(args.warmup)
    rk.sync()
    # the dominant kernel: event time per step of each candidate over two untimed steps
    per_kernel = {}
    for name in CANDIDATES:
        ms, n, _ = profile(rk, name, 2)
        per_kernel[name] = ms / 2
    dominant = max(per_kernel, key=per_kernel.get)
    log("kernel ms/step:", json.dumps({k: round(v, 4) for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1])}))

    # timed region: K graph-replayed steps, bracketed by device synchronisations
    rk.sync()
    t0 = time.perf_counter()
    rk.step(args.steps)
    rk.sync()
    t1 = time.perf_counter()
    # the dominant kernel over K more steps, its launches bracketed by events on the domain's stream
    ev_ms, nl, b = profile(rk, dominant, args.steps)
    bad = rk.check_finite()
    if bad != -1:
        raise SystemExit(f"non-finite values in field {bad}")
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = g.points * args.steps / (t1 - t0)
    avg_launch_s = ev_ms / 1e3 / nl
    achieved = b / nl / avg_launch_s / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "dry warm bubble 64x64x40, 1 km, 20 km top, plain FP64, bit-identical to the "
                               "reference CPU FP64 build with scalar glibc pow",
                   "nx": g.nx, "ny": g.ny, "nz": g.nz, "dx_m": g.dx, "dt_s": g.dt, "time_step_sound": rk.n_sound,
                   "damp_opt": 3, "h_adv_order": 5, "v_adv_order": 3, "precision": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": b / nl, "avg_launch_us": avg_launch_s * 1e6,
                     "launches": nl},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, prm)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
