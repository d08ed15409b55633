#!/usr/bin/env python3
"""Throughput of the RK3 split-explicit dynamics step on one MI355X, FP32-bulk against plain FP64.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--workload hills|bubble64]

Workload "hills" (default): the real-data-shaped 119x89x44 domain at 15 km (seeded hills up to 2.5 km high
and 30-150 km wide, US standard atmosphere, seeded u/v jets of 5-40 m/s with +-1 m/s noise, map factors
in [0.95, 1.05], Coriolis, a hybrid coordinate with a 50 hPa top, six moist species advected with the
positive-definite limiter, qv from 80 % relative humidity falling to 10 % at 10 km), dt 90 s with 4
acoustic steps, Rayleigh damping of w in the top 5 km.  Its lateral boundaries are specified: a 5-point
zone (1 spec + 4 relaxation rows) nudged toward boundary data that start at the initial state and drift
with a seeded 6-hour oscillation of the winds, theta and qv (cppwrf_amd.boundary.drifting_boundary); the
other moist species take flow-dependent boundary values.
Workload "bubble64": the 64x64x40 dry warm bubble at 1 km (dt 6 s, 6 acoustic steps, damp_opt 3).

A step is one RK3 large step (three stages and their acoustic substeps, scalar transport included).  The
timed region replays the step's HIP graph K times with the state resident in HBM, bracketed by device
synchronisations; it is measured for the FP64 domain and for the FP32-bulk domain of the same initial
state.  `value` is the FP32-bulk throughput in grid-point-steps/s (nx*ny*nz*steps/s); the FP64 figure,
the speed-up and the relative L2 / max error of every variable of the FP32-bulk state against the FP64
state after the same steps are printed beside it.

The roofline is that of the dominant kernel of the FP32-bulk step: HIP events recorded on the domain's
stream around each of its launches during K further steps run launch by launch (events cannot be read
back from inside a replayed graph); algorithmic bytes = distinct 3-D arrays it reads and writes x their
storage size x the points of one launch.  The CPU baseline is the C++/OpenMP oracle (the FP64 restatement
of the reference) on this host's cores, a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "tests"))

METRIC = "grid-point-steps/s FP32-batched vs FP64 plain at stated L2; p50/p99 forecast latency"
HBM_PEAK_GBS = 8000.0
# kernels whose event timing is sampled to find the dominant one
CANDIDATES = ["large_prep", "cq", "tend_u", "tend_v", "tend_w", "tend_t", "tend_ph", "ss_prep", "ss_column", "uv",
              "acoustic", "finish", "sc_update", "sc_pd_pre", "sc_pd_scale", "sc_pd_update", "sixth", "sixth_sc",
              "set_w_surface"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    """(case, parameters, description, boundary data or None)"""
    from cppwrf_amd import _abi as A
    from cppwrf_amd.boundary import drifting_boundary
    from cppwrf_amd.dyn import Case
    if name == "bubble64":
        case = Case.bubble3d(64, 64, 40, dx=1000.0, ztop=20000.0, dt=6.0, seed=1, noise_amp=0.05)
        kw = dict(time_step_sound=6, damp_opt=3, zdamp=5000.0, dampcoef=0.2)
        desc = {"workload": "dry warm bubble 64x64x40, 1 km, 20 km top, periodic, damp_opt 3, dt 6 s",
                "nx": 64, "ny": 64, "nz": 40, "dx_m": 1000.0, "dt_s": 6.0, "time_step_sound": 6}
        bdy = None
    else:
        case = Case.hills(119, 89, 44, dx=15000.0, dt=90.0, seed=1, bc=A.MI_BC_SPECIFIED)
        kw = dict(time_step_sound=4, damp_opt=3, zdamp=5000.0, dampcoef=0.2, n_moist=6, n_scalars=6,
                  moist_adv_opt=1)
        desc = {"workload": "real-data-shaped 119x89x44 at 15 km: 12 seeded hills, US standard atmosphere, jets, "
                            "map factors, Coriolis, hybrid coordinate (50 hPa top), 6 moist species with the PD "
                            "limiter, dt 90 s, 4 acoustic steps; specified lateral boundaries (5-point relaxation "
                            "zone, seeded slowly varying boundary data)",
                "nx": 119, "ny": 89, "nz": 44, "dx_m": 15000.0, "dt_s": 90.0, "time_step_sound": 4,
                "moist_species": 6, "lateral_bc": "specified"}
        bdy = drifting_boundary(case, seed=17)
    return case, kw, desc, bdy


def measured_traffic(workload_name, kernel, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
    workload (each counter scaled by its reading for a copy kernel of known size), or None.  The first line of a
    kernel and precision is its most-dispatched block shape, the autotuned one."""
    path = os.path.join(HERE, "profiles", "hbm_traffic_hills119x89x44.md")
    if workload_name != "hills" or not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip()
        if not line.startswith("traffic:"):
            continue
        kv = dict(x.split("=", 1) for x in line.split()[1:])
        if kv.get("kernel") == kernel and kv.get("precision") == precision:
            return {"bytes": float(kv["bytes"]), "shape": kv.get("shape"),
                    "source": "profiles/hbm_traffic_hills119x89x44.md"}
    return None


def profile(rk, name, steps):
    from cppwrf_amd._abi import check, lib
    check(lib().mi_profile_kernel(rk.handle, name.encode()))
    rk.step(steps)
    ms, n, b = C.c_double(), C.c_int(), C.c_double()
    check(lib().mi_profile_read(rk.handle, C.byref(ms), C.byref(n), C.byref(b)))
    check(lib().mi_profile_kernel(rk.handle, None))
    return ms.value, n.value, b.value


def timed_steps(rk, steps):
    rk.sync()
    t0 = time.perf_counter()
    rk.step(steps)
    rk.sync()
    return time.perf_counter() - t0


def error_report(d32, d64, n_moist):
    from cppwrf_amd import _abi as A
    ids = {"U": A.MI_U, "V": A.MI_V, "W": A.MI_W, "T": A.MI_T, "PH": A.MI_PH, "MU": A.MI_MU}
    if n_moist > 0:
        ids["QVAPOR"] = A.MI_Q
    out = {}
    for name, fid in ids.items():
        a, b = d32.field(fid), d64.field(fid)
        nb = float(np.linalg.norm(b))
        out[name] = {"rel_l2": float(np.linalg.norm(a - b)) / nb if nb > 0 else 0.0, "max_abs": float(np.abs(a - b).max())}
    return out


def ensemble_bench(ctx, members=32, steps=10, seed=2024):
    """The batched ensemble: `members` domains of uneven sizes (nx 60-220, ny 60-180, nz 35/44/50, dx 3-15 km),
    each the 3-D warm bubble with its own seed plus +-0.5 K theta and +-1 m/s wind perturbations, dt = 6 s per
    km of dx, 6 acoustic steps, damp_opt 3; all members advance in one HIP graph per step (mi_batch).  The same
    ensemble is stepped in FP64 and in FP32-bulk; the FP32 error is the largest relative L2 over the members."""
    from cppwrf_amd import _abi as A
    from cppwrf_amd.dyn import Batch, Case, RK3
    rng = np.random.default_rng(seed)
    specs = []
    for m in range(members):
        nx, ny = int(rng.integers(60, 221)), int(rng.integers(60, 181))
        nz, dx = int(rng.choice([35, 44, 50])), float(rng.uniform(3000.0, 15000.0))
        specs.append((nx, ny, nz, dx, int(rng.integers(1 << 30))))
    pts = sum(nx * ny * nz for nx, ny, nz, _, _ in specs)
    out, states = {"members": members, "steps": steps, "grid_points": pts}, {}
    for tag, prec in (("fp64", A.MI_FP64), ("fp32", A.MI_FP32_BULK)):
        doms = []
        for nx, ny, nz, dx, sd in specs:
            case = Case.bubble3d(nx, ny, nz, dx=dx, ztop=20000.0, dt=6.0 * dx / 1000.0, seed=sd, noise_amp=0.05)
            rk = RK3(ctx, case.grid, case.params(time_step_sound=6, damp_opt=3, zdamp=5000.0, precision=prec), case)
            prng = np.random.default_rng(sd)
            for fid, amp in ((A.MI_T, 0.5), (A.MI_U, 1.0), (A.MI_V, 1.0)):
                f = case.field(fid)
                rk.set_field(fid, f + prng.uniform(-amp, amp, size=f.shape))
            doms.append(rk)
        b = Batch(ctx, doms)
        b.step(1)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        dt = time.perf_counter() - t0
        out[tag] = {"value": pts * steps / dt, "unit": "grid-point-steps/s", "ms_per_step": dt * 1e3 / steps}
        states[tag] = [{fid: d.field(fid) for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T)} for d in doms]
        b.close()
        for d in doms:
            d.close()
    worst = {}
    for name, fid in (("U", A.MI_U), ("V", A.MI_V), ("W", A.MI_W), ("T", A.MI_T)):
        worst[name] = max(float(np.linalg.norm(a[fid] - b[fid]) / np.linalg.norm(b[fid]))
                          for a, b in zip(states["fp32"], states["fp64"]))
    out["speedup_fp32_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    out["fp32_max_rel_l2_over_members"] = worst
    return out


def serving_bench(ctx, seconds=20.0, rate=2.0, scale=0.5, seed=1):
    """The serving trace (Poisson arrivals, domains 50-300 x 50-300 x 35/44/50 at 3-15 km, forecasts of 1-6 h,
    the warm-bubble recipe per request) replayed in real time through the continuous-batching server in
    FP32-bulk, horizontal sizes and forecast lengths scaled by `scale`; 5 % of the requests (at least one) are
    rerun alone in FP64 for the error of the served FP32-bulk results."""
    from cppwrf_amd import _abi as A
    from cppwrf_amd.dyn import RK3
    from cppwrf_amd.serving import poisson_trace, run_trace
    reqs = poisson_trace(seconds, rate, seed, scale, A.MI_FP32_BULK)
    keep = {r.rid for r in reqs[:: max(1, len(reqs) // max(1, len(reqs) // 20))]}
    res = run_trace(reqs, 200 << 30, ctx=ctx, keep=keep)
    errs = {}
    for r in reqs:
        if r.result is None:
            continue
        r.precision = A.MI_FP64
        case = r.case()
        d = RK3(ctx, case.grid, r.params(case), case)
        d.step(r.steps)
        for name, fid in (("U", A.MI_U), ("V", A.MI_V), ("W", A.MI_W), ("T", A.MI_T), ("MU", A.MI_MU)):
            b = d.field(fid)
            e = float(np.linalg.norm(r.result[fid] - b) / max(np.linalg.norm(b), 1e-300))
            errs[name] = max(errs.get(name, 0.0), e)
        d.close()
    res.update({"trace_seconds": seconds, "rate_per_s": rate, "scale": scale, "precision": "fp32-bulk",
                "fp64_reruns": len(keep), "fp32_max_rel_l2_vs_fp64_reruns": errs})
    return res


def cpu_baseline(case, prm, bdy, budget_s=15.0):
    """The oracle (C++/OpenMP, bit-identical to the reference CPU FP64 build with scalar pow) on this host's cores."""
    from oracle_lib import OracleDomain, olib
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    olib().orc_set_threads(threads)
    orc = OracleDomain(case, prm)
    if bdy:
        orc.load_boundary(bdy)
    t = time.perf_counter()
    orc.step(1)
    one = time.perf_counter() - t
    n = max(1, min(300, int(budget_s / max(one, 1e-3))))
    t = time.perf_counter()
    orc.step(n)
    dt = time.perf_counter() - t
    g = case.grid
    return {"value": g.nx * g.ny * g.nz * n / dt, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (FP64), {n} steps of the same {g.nx}x{g.ny}x{g.nz} workload after 1 untimed step, "
                      f"{threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["hills", "bubble64"], default="hills")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default block shapes")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 32-member batched ensemble")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving trace")
    ap.add_argument("--error-hours", type=int, default=6,
                    help="forecast hours over which the FP32-bulk error is reported hourly (0: only after the timed steps)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (each forecast domain runs on a "
              f"single GPU; independent domains shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)

    from cppwrf_amd import _abi as A
    from cppwrf_amd.dyn import Context, RK3

    case, kw, desc, bdy = workload(args.workload)
    g = case.grid
    ctx = Context(0)
    doms = {}
    for tag, prec in (("fp64", A.MI_FP64), ("fp32", A.MI_FP32_BULK)):
        rk = RK3(ctx, g, case.params(precision=prec, **kw), case)
        if bdy:
            rk.load_boundary(bdy)
        if not args.no_autotune:
            log(f"{tag}: autotuned block shapes", rk.autotune(1))  # steps the domain, then restores its state
        doms[tag] = rk
    # the same steps on both: warm-up, then the timed region
    secs = {}
    for tag, rk in doms.items():
        rk.step(args.warmup)
        secs[tag] = timed_steps(rk, args.steps)
        log(f"{tag}: {secs[tag] * 1e3 / args.steps:.3f} ms/step")
    for tag, rk in doms.items():
        bad = rk.check_finite()
        if bad != -1:
            raise SystemExit(f"{tag}: non-finite values in field {bad}")
    err = error_report(doms["fp32"], doms["fp64"], kw.get("n_moist", 0))
    # the FP32-bulk error per forecast hour up to 6 h: both domains stepped on to each whole hour
    done = args.warmup + args.steps
    hourly = []
    for hour in range(1, args.error_hours + 1):
        target = int(np.ceil(hour * 3600.0 / g.dt - 1e-9))
        if target > done:
            for rk in doms.values():
                rk.step(target - done)
            done = target
        if target == done:
            hourly.append({"hour": hour, "steps": done,
                           "per_variable": error_report(doms["fp32"], doms["fp64"], kw.get("n_moist", 0))})
    pts = g.nx * g.ny * g.nz
    v64 = pts * args.steps / secs["fp64"]
    v32 = pts * args.steps / secs["fp32"]

    # the dominant kernel of the FP32-bulk step: event time per step of each candidate over two steps
    rk = doms["fp32"]
    per_kernel = {}
    for name in CANDIDATES:
        ms, n, _ = profile(rk, name, 2)
        if n:
            per_kernel[name] = ms / 2
    dominant = max(per_kernel, key=per_kernel.get)
    log("fp32 kernel ms/step:", json.dumps({k: round(v, 4) for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1])}))
    ev_ms, nl, b = profile(rk, dominant, args.steps)
    avg_launch_s = ev_ms / 1e3 / nl
    achieved = b / nl / avg_launch_s / 1e9
    traffic = measured_traffic(args.workload, dominant, "fp32")

    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs["fp32"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-bulk (f64 islands)",
        "data": "synthetic",
        "config": dict(desc, precision="fp32-bulk"),
        "fp64": {"value": v64, "ms_per_step": secs["fp64"] * 1e3 / args.steps, "dtype": "f64",
                 "note": "plain FP64, bit-identical to the reference CPU FP64 build with CPPWRF_NO_LIBMVEC (scalar glibc pow in calc_p_rho_phi)"},
        "speedup_fp32_vs_fp64": v32 / v64,
        "fp32_error_vs_fp64": {"after_steps": args.warmup + args.steps,
                               "simulated_s": (args.warmup + args.steps) * g.dt, "per_variable": err,
                               "hourly": hourly,
                               "max_rel_l2_6h": max(v["rel_l2"] for v in hourly[-1]["per_variable"].values())
                               if hourly and hourly[-1]["hour"] == 6 else None},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic, "alg_bytes_per_launch": b / nl,
                     "avg_launch_us": avg_launch_s * 1e6, "launches": nl, "precision": "fp32-bulk"},
    }
    for d in doms.values():
        d.close()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(ctx)
        log("ensemble:", json.dumps(out["ensemble"]))
    if not args.no_serving:
        out["serving"] = serving_bench(ctx)
        log("serving:", json.dumps(out["serving"]))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, case.params(**kw), bdy)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
