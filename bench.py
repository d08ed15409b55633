#!/usr/bin/env python3
"""Throughput of the RK3 split-explicit dynamics step on one MI355X, FP32-bulk batched against plain FP64.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

`value` (the headline) is the batched ensemble in FP32-bulk: 32 members of uneven sizes (nx 60-220, ny 60-180,
nz 35/44/50, dx 3-15 km; the 3-D warm-bubble recipe with per-member seeds and perturbations, dt 6 s per km of dx,
6 acoustic steps, damp_opt 3) advanced together, one HIP graph per batch step, in grid-point-steps/s
(sum of nx*ny*nz over the members x steps / s).  The same ensemble in plain FP64 is timed beside it, with the
speed-up and the largest per-member relative L2 error of FP32-bulk against FP64 after the same steps.

Beside it (cppwrf_amd/workloads.py):
 - "c3": the 425x300x50 real-data-shaped domain at 3 km (terrain, hybrid coordinate, map factors, six moist
   species with the positive-definite limiter, sixth-order diffusion, specified boundaries), FP32-bulk and FP64;
   the roofline of the FP32-bulk step's dominant kernel is measured here (its working set is far beyond the
   256 MB Infinity Cache): HIP events on the domain's stream around each of its launches during K further steps
   run launch by launch (events cannot be read back from inside a replayed graph), algorithmic bytes = distinct
   3-D arrays it reads and writes x their storage size x the points of one launch;
 - "c2": the 119x89x44 domain at 15 km, FP32-bulk and FP64, with the FP32-bulk error per forecast hour to 6 h;
 - "serving": a short Poisson trace through the continuous-batching server;
 - "cpu_baseline": the C++/OpenMP oracle (FP64, bit-identical to the reference CPU build with scalar pow) on this
   host's cores, a bounded sample of the 119x89x44 workload.

A step is one RK3 large step (three stages, their acoustic substeps, scalar transport).  Timed regions replay
HIP graphs with the state resident in HBM, bracketed by device synchronisations.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

# the CPU baseline's OpenMP threads pinned one per core, close together: set before any OpenMP runtime loads
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

HERE = os.path.dirname(os.path.abspath(__file__))
# the CPUs this process may run on, read before OpenMP binds the main thread to one of them
AFFINITY_CPUS = len(os.sched_getaffinity(0))
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


def measured_traffic(profile_name, kernel, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of a workload
    (each counter scaled by its reading for a copy kernel of known size), or None.  The first line of a kernel and
    precision is its most-dispatched block shape, the autotuned one."""
    path = os.path.join(HERE, "profiles", profile_name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip()
        if not line.startswith("traffic:"):
            continue
        kv = dict(x.split("=", 1) for x in line.split()[1:])
        if kv.get("kernel") == kernel and kv.get("precision") == precision:
            return {"bytes": float(kv["bytes"]), "shape": kv.get("shape"), "source": "profiles/" + profile_name}
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


def ensemble_bench(ctx, steps, warmup, autotune):
    """the batched ensemble in FP64 and FP32-bulk: timing and the per-member FP32-bulk error"""
    from cppwrf_amd import _abi as A
    from cppwrf_amd import workloads as W
    from cppwrf_amd.dyn import Batch
    specs = W.ensemble_specs()
    pts = sum(nx * ny * nz for nx, ny, nz, _, _ in specs)
    out, doms = {"members": len(specs), "grid_points": pts}, {}
    for tag, prec in (("fp64", A.MI_FP64), ("fp32", A.MI_FP32_BULK)):
        ms = [W.ensemble_member(ctx, sp, prec) for sp in specs]
        if autotune:
            for d in ms:
                d.autotune(1)
        # both batch schedules (the same bits per member): a branch per member, and the fused substep launches
        b = Batch(ctx, ms)
        per = {}
        for sched, sname in ((A.MI_BATCH_BRANCHES, "branches"), (A.MI_BATCH_FUSED, "fused")):
            b.set_schedule(sched)
            b.step(warmup)
            b.sync()
            t0 = time.perf_counter()
            b.step(steps)
            b.sync()
            per[sname] = time.perf_counter() - t0
            log(f"ensemble {tag} {sname}: {per[sname] * 1e3 / steps:.3f} ms/step, "
                f"{pts * steps / per[sname]:.3e} grid-point-steps/s")
        # the faster schedule times the headline's K steps
        best = min(per, key=per.get)
        b.set_schedule(A.MI_BATCH_BRANCHES if best == "branches" else A.MI_BATCH_FUSED)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        dt = time.perf_counter() - t0
        log(f"ensemble {tag}: {best} schedule, {dt * 1e3 / steps:.3f} ms/step")
        out[tag] = {"value": pts * steps / dt, "unit": "grid-point-steps/s", "ms_per_step": dt * 1e3 / steps,
                    "schedule": best, "trial_by_schedule": {k: pts * steps / v for k, v in per.items()}}
        b.close()
        doms[tag] = ms
    worst = {}
    for d32, d64 in zip(doms["fp32"], doms["fp64"]):
        for name, v in W.error_report(d32, d64, False).items():
            worst[name] = max(worst.get(name, 0.0), v["rel_l2"])
    for ms in doms.values():
        for d in ms:
            d.close()
    out["speedup_fp32_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    out["fp32_max_rel_l2_over_members"] = worst
    out["after_steps"] = 2 * (warmup + steps) + steps
    return out


def ensemble_error_hours(ctx, hours):
    """every member of the ensemble from its initial state in FP32-bulk and in FP64, side by side for `hours`
    forecast hours (each member its own number of steps per hour, dt 6 s per km): the largest relative L2 over the
    members per variable at each hour"""
    from cppwrf_amd import _abi as A
    from cppwrf_amd import workloads as W
    worst = [dict() for _ in range(hours)]
    steps = 0
    for q, sp in enumerate(W.ensemble_specs()):
        dt = 6.0 * sp[3] / 1000.0
        d32, d64 = W.ensemble_member(ctx, sp, A.MI_FP32_BULK), W.ensemble_member(ctx, sp, A.MI_FP64)
        done = 0
        for h in range(hours):
            target = int(round((h + 1) * 3600.0 / dt))
            d32.step(target - done)
            d64.step(target - done)
            done = target
            for name, v in W.error_report(d32, d64, False).items():
                worst[h][name] = max(worst[h].get(name, 0.0), v["rel_l2"])
        if d32.check_finite() != -1 or d64.check_finite() != -1:
            raise RuntimeError(f"ensemble member {q} is not finite after {hours} h")
        steps += done
        d32.close()
        d64.close()
        log(f"ensemble member {q}: {done} steps to {hours} h, worst so far at {hours} h:",
            {k: f"{v:.2e}" for k, v in worst[-1].items()})
    return {"hours": hours, "member_steps": steps,
            "hourly_max_rel_l2_over_members": [{"hour": h + 1, "per_variable": worst[h]} for h in range(hours)],
            "max_rel_l2_last_hour": max(worst[-1].values())}


def domain_bench(ctx, name, steps, warmup, autotune, error_hours):
    """one real-data-shaped domain (c2 or c3) in FP64 and FP32-bulk: timing, the FP32-bulk error after the timed
    steps and per forecast hour to `error_hours`, and the FP32-bulk domain (for the roofline)"""
    from cppwrf_amd import _abi as A
    from cppwrf_amd import workloads as W
    case, kw, desc, bdy, extra = W.hills_c2() if name == "c2" else W.hills_c3()
    g = case.grid
    pts = g.nx * g.ny * g.nz
    doms, secs = {}, {}
    for tag, prec in (("fp64", A.MI_FP64), ("fp32", A.MI_FP32_BULK)):
        rk = W.make_domain(ctx, case, kw, bdy, extra, prec)
        if autotune:
            log(f"{name} {tag}: autotuned block shapes", rk.autotune(1))  # restores the state afterwards
        rk.step(warmup)
        secs[tag] = timed_steps(rk, steps)
        log(f"{name} {tag}: {secs[tag] * 1e3 / steps:.3f} ms/step")
        doms[tag] = rk
    for tag, rk in doms.items():
        if rk.check_finite() != -1:
            raise SystemExit(f"{name} {tag}: non-finite values")
    out = dict(desc)
    out["fp32"] = {"value": pts * steps / secs["fp32"], "ms_per_step": secs["fp32"] * 1e3 / steps}
    out["fp64"] = {"value": pts * steps / secs["fp64"], "ms_per_step": secs["fp64"] * 1e3 / steps}
    out["speedup_fp32_vs_fp64"] = secs["fp64"] / secs["fp32"]
    done = warmup + steps
    out["fp32_error_vs_fp64"] = {"after_steps": done, "per_variable": W.error_report(doms["fp32"], doms["fp64"], True)}
    hourly = []
    for hour in range(1, error_hours + 1):
        target = int(np.ceil(hour * 3600.0 / g.dt - 1e-9))
        if target > done:
            for rk in doms.values():
                rk.step(target - done)
            done = target
        if target == done:
            hourly.append({"hour": hour, "steps": done, "per_variable": W.error_report(doms["fp32"], doms["fp64"], True)})
    if hourly:
        out["fp32_error_vs_fp64"]["hourly"] = hourly
        out["fp32_error_vs_fp64"]["max_rel_l2_last_hour"] = max(v["rel_l2"] for v in hourly[-1]["per_variable"].values())
    doms["fp64"].close()
    return out, doms["fp32"]


def roofline(rk, steps, profile_name, workload="c3 (425x300x50)"):
    """the dominant kernel of the FP32-bulk step: event time per step of each candidate over two steps, then its
    average launch over `steps` steps against the algorithmic bytes of a launch"""
    per_kernel = {}
    for name in CANDIDATES:
        ms, n, _ = profile(rk, name, 2)
        if n:
            per_kernel[name] = ms / 2
    dominant = max(per_kernel, key=per_kernel.get)
    log("fp32 kernel ms/step:", json.dumps({k: round(v, 4) for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1])}))
    ev_ms, nl, b = profile(rk, dominant, steps)
    avg_launch_s = ev_ms / 1e3 / nl
    achieved = b / nl / avg_launch_s / 1e9
    traffic = measured_traffic(profile_name, dominant, "fp32")
    return {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
            "traffic_source": traffic, "alg_bytes_per_launch": b / nl, "avg_launch_us": avg_launch_s * 1e6,
            "launches": nl, "precision": "fp32-bulk", "workload": workload,
            "kernel_ms_per_step": {k: round(v, 4) for k, v in per_kernel.items()}}


def serving_bench(ctx, seconds, rate=2.0, scale=1.0, seed=1):
    """the serving trace (Poisson arrivals, domains 50-300 x 50-300 x 35/44/50 at 3-15 km, forecasts of 1-6 h,
    the warm-bubble recipe per request) replayed in real time through the continuous-batching server in
    FP32-bulk; 5 % of the requests (at least one) are rerun alone in FP64 for the error of the served results"""
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


def physical_cores():
    """distinct (package, core) pairs of /proc/cpuinfo, or None"""
    try:
        cores, phys = set(), None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":")[1].strip()
            elif line.startswith("core id"):
                cores.add((phys, line.split(":")[1].strip()))
        return len(cores) or None
    except OSError:
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(budget_s=15.0):
    """The oracle (C++/OpenMP, bit-identical to the reference CPU FP64 build with scalar pow) on this host's cores,
    a bounded sample of the 119x89x44 workload."""
    from cppwrf_amd import workloads as W
    from oracle_lib import OracleDomain, olib
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    olib().orc_set_threads(threads)
    case, kw, _, bdy, extra = W.hills_c2()
    orc = OracleDomain(case, case.params(**kw))
    orc.load_boundary(bdy)
    for fid, a in extra.items():
        orc.set_field(fid, a)
    t = time.perf_counter()
    orc.step(1)
    one = time.perf_counter() - t
    n = max(1, min(300, int(budget_s / max(one, 1e-3))))
    t = time.perf_counter()
    orc.step(n)
    dt = time.perf_counter() - t
    g = case.grid
    return {"value": g.nx * g.ny * g.nz * n / dt, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "physical_cores": physical_cores(), "affinity_cpus": AFFINITY_CPUS,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
            "cpu_model": cpu_model(), "oracle_flags": "-O3 -march=x86-64-v3 -ffp-contract=off -fopenmp",
            "sample": f"oracle (FP64), {n} steps of the 119x89x44 real-data-shaped workload after 1 untimed step, "
                      f"{threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default block shapes")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched ensemble (for profiling runs: the headline then falls back to c3)")
    ap.add_argument("--no-c2", action="store_true", help="skip the 119x89x44 domain")
    ap.add_argument("--domain", choices=("c3", "c2"), default="c3",
                    help="with --no-ensemble: the single domain to time (c3 425x300x50, c2 119x89x44)")
    ap.add_argument("--no-c3", action="store_true", help="skip the 425x300x50 domain (and the roofline)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving trace")
    ap.add_argument("--serving-seconds", type=float, default=20.0, help="length of the serving trace")
    ap.add_argument("--error-hours", type=int, default=6,
                    help="forecast hours over which the 119x89x44 FP32-bulk error is reported hourly")
    ap.add_argument("--ensemble-hours", type=int, default=6,
                    help="forecast hours over which every ensemble member's FP32-bulk error is reported hourly "
                         "(0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (each forecast domain runs on a "
              f"single GPU; independent domains shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)

    from cppwrf_amd.dyn import Context
    ctx = Context(0)
    if args.no_ensemble:
        name = args.domain
        grid = "425x300x50" if name == "c3" else "119x89x44"
        one, d32 = domain_bench(ctx, name, args.steps, args.warmup, not args.no_autotune,
                                0 if name == "c3" else args.error_hours)
        out = {"metric": METRIC, "value": one["fp32"]["value"], "unit": "grid-point-steps/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": one["fp32"]["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32-bulk (f64 islands)",
               "data": "synthetic", "config": {"workload": f"{grid} single domain (the ensemble was skipped)"},
               "roofline": roofline(d32, args.steps, f"hbm_traffic_hills{grid}.md", f"{name} ({grid})"), name: one}
        d32.close()
        print(json.dumps(out), flush=True)
        return
    ens = ensemble_bench(ctx, args.steps, args.warmup, not args.no_autotune)
    out = {
        "metric": METRIC,
        "value": ens["fp32"]["value"],
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ens["fp32"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-bulk (f64 islands)",
        "data": "synthetic",
        "config": {"workload": "batched ensemble: 32 members of uneven sizes (nx 60-220, ny 60-180, nz 35/44/50, "
                               "dx 3-15 km), 3-D warm bubble per member seed with +-0.5 K / +-1 m/s perturbations, "
                               "dt 6 s per km, 6 acoustic steps, damp_opt 3, one HIP graph per batch step",
                   "members": ens["members"], "grid_points": ens["grid_points"], "precision": "fp32-bulk",
                   "batch_schedule": ens["fp32"]["schedule"], "trial_by_schedule": ens["fp32"]["trial_by_schedule"]},
        "fp64": {"value": ens["fp64"]["value"], "ms_per_step": ens["fp64"]["ms_per_step"], "dtype": "f64",
                 "batch_schedule": ens["fp64"]["schedule"], "trial_by_schedule": ens["fp64"]["trial_by_schedule"],
                 "note": "plain FP64, bit-identical to the reference CPU FP64 build with CPPWRF_NO_LIBMVEC (scalar "
                         "glibc pow in calc_p_rho_phi)"},
        "speedup_fp32_vs_fp64": ens["speedup_fp32_vs_fp64"],
        "fp32_error_vs_fp64": {"after_steps": ens["after_steps"],
                               "max_rel_l2_over_members": ens["fp32_max_rel_l2_over_members"]},
    }
    if args.ensemble_hours > 0:
        out["fp32_error_vs_fp64"]["forecast"] = ensemble_error_hours(ctx, args.ensemble_hours)
    if not args.no_c3:
        c3, d32 = domain_bench(ctx, "c3", args.steps, args.warmup, not args.no_autotune, 0)
        out["roofline"] = roofline(d32, args.steps, "hbm_traffic_hills425x300x50.md")
        d32.close()
        out["c3"] = c3
    if not args.no_c2:
        c2, d32 = domain_bench(ctx, "c2", max(args.steps, 20), args.warmup, not args.no_autotune, args.error_hours)
        d32.close()
        out["c2"] = c2
    if not args.no_serving:
        out["serving"] = serving_bench(ctx, args.serving_seconds)
        log("serving:", json.dumps(out["serving"]))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
