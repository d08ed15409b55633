"""FP32-bulk against plain FP64 over a forecast, hour by hour (the error metric of the reference's FP32 study,
docs/FP32_ACCURACY.md:24-35: relative L2 ||a - b|| / ||b|| and max |a - b| per variable, b the FP64 run).

    python -m cppwrf_amd.accuracy [--size 425x300x50] [--hours 6]

The default is the 425x300x50 real-data-shaped domain at 3 km (seeded hills of 10-60 km width, dt 18 s, 6 acoustic
steps, six moist species with the positive-definite limiter, sixth-order diffusion with factor 0.12, periodic
boundaries): 200 steps per forecast hour.  Prints one JSON line per hour and a summary line."""
import argparse
import json
import time

import numpy as np

from . import _abi as A
from .dyn import Case, Context, RK3

VARS = {"U": A.MI_U, "V": A.MI_V, "W": A.MI_W, "T": A.MI_T, "PH": A.MI_PH, "MU": A.MI_MU, "QVAPOR": A.MI_Q}


def errors(d32, d64):
    out = {}
    for name, fid in VARS.items():
        a, b = d32.field(fid), d64.field(fid)
        nb = float(np.linalg.norm(b))
        out[name] = {"rel_l2": float(np.linalg.norm(a - b)) / nb if nb > 0 else 0.0,
                     "max_abs": float(np.abs(a - b).max())}
    return out


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--size", default="425x300x50")
    ap.add_argument("--hours", type=int, default=6)
    ap.add_argument("--dx", type=float, default=3000.0)
    ap.add_argument("--dt", type=float, default=18.0)
    args = ap.parse_args()
    nx, ny, nz = (int(v) for v in args.size.split("x"))
    case = Case.hills(nx, ny, nz, dx=args.dx, dt=args.dt, seed=1, hill_wmin=10000.0, hill_wmax=60000.0)
    kw = dict(time_step_sound=6, damp_opt=3, zdamp=5000.0, dampcoef=0.2, n_moist=6, n_scalars=6, moist_adv_opt=1,
              diff_6th_opt=1, diff_6th_factor=0.12)
    ctx = Context(0)
    d64 = RK3(ctx, case.grid, case.params(precision=A.MI_FP64, **kw), case)
    d32 = RK3(ctx, case.grid, case.params(precision=A.MI_FP32_BULK, **kw), case)
    per_hour = int(round(3600.0 / args.dt))
    worst = {}
    t0 = time.perf_counter()
    for hour in range(1, args.hours + 1):
        d64.step(per_hour)
        d32.step(per_hour)
        if d32.check_finite() != -1 or d64.check_finite() != -1:
            raise SystemExit(f"non-finite values after hour {hour}")
        e = errors(d32, d64)
        for k, v in e.items():
            worst[k] = max(worst.get(k, 0.0), v["rel_l2"])
        print(json.dumps({"hour": hour, "steps": hour * per_hour, "per_variable": e}), flush=True)
    print(json.dumps({"size": args.size, "dx_m": args.dx, "dt_s": args.dt, "hours": args.hours,
                      "max_rel_l2": worst, "max_over_variables": max(worst.values()),
                      "wall_s": time.perf_counter() - t0}), flush=True)


if __name__ == "__main__":
    main()
