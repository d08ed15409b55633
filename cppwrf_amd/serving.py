"""Continuous-batching forecast server on one GPU.

Requests arrive over time, each one forecast: a domain (its own size, resolution and options), an initial
state and a length in steps.  The server keeps every admitted forecast's state resident on the device and
advances all of them together, one batch step per tick (mi_batch: one HIP graph holding every member's step
as a branch, so the members' kernels overlap).  At each step boundary it retires the forecasts that are
done, admits waiting ones while the HBM budget allows, and re-forms the batch.  It reports time-to-forecast
(arrival to completion) percentiles and the aggregate throughput in grid-point-steps/s.

    python -m cppwrf_amd.serving --seconds 30 --rate 2 [--precision fp32|fp64]

prints one JSON line: p50/p99 time-to-forecast, throughput, admission and memory statistics.  The trace is
the synthetic Poisson one of the serving workload: nx, ny uniform in [50, 300], nz in {35, 44, 50}, dx
uniform in [3, 15] km, dt = 6 s per km of dx, forecast length uniform in [1, 6] h, each request the 3-D
warm-bubble recipe with its own seed (lengths and sizes scale with --scale for short runs).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from .dyn import Batch, Case, Context, RK3


@dataclass
class Request:
    rid: int
    nx: int
    ny: int
    nz: int
    dx: float
    steps: int
    seed: int
    arrival: float = 0.0     # s since the start of the trace
    precision: int = A.MI_FP64
    n_sound: int = 6
    # filled in by the server
    admitted: float = math.nan
    done: float = math.nan
    domain: RK3 | None = None
    taken: int = 0
    result: dict | None = None

    @property
    def points(self) -> int:
        return self.nx * self.ny * self.nz

    def case(self):
        dt = 6.0 * self.dx / 1000.0
        return Case.bubble3d(self.nx, self.ny, self.nz, dx=self.dx, ztop=20000.0, dt=dt, seed=self.seed,
                             noise_amp=0.05, bubble_zc=2000.0)

    def params(self, case):
        return case.params(time_step_sound=self.n_sound, damp_opt=3, zdamp=5000.0, dampcoef=0.2, precision=self.precision)


def device_bytes_estimate(r: Request) -> int:
    """Device memory of a domain before it exists: the padded allocation of the library's layout
    (row pitch a multiple of 32 elements, a halo of 3) times the number of 3-D arrays of a dry domain."""
    pitch = (8 + r.nx + 1 + 3 + 31) // 32 * 32
    rows = r.ny + 1 + 6
    per3 = pitch * rows * (r.nz + 1) * (4 if r.precision == A.MI_FP32_BULK else 8)
    return int(per3 * 52 + pitch * rows * 8 * 30)


@dataclass
class ServerStats:
    admitted: int = 0
    completed: int = 0
    ticks: int = 0
    point_steps: float = 0.0
    peak_bytes: int = 0
    peak_members: int = 0
    latencies: list = field(default_factory=list)


class ForecastServer:
    def __init__(self, ctx: Context, hbm_budget: int, max_members: int = 64, clock=time.perf_counter):
        self.ctx = ctx
        self.clock = clock
        self.budget = hbm_budget
        self.max_members = max_members
        self.waiting: list[Request] = []
        self.active: list[Request] = []
        self.finished: list[Request] = []
        self.batch: Batch | None = None
        self.bytes_in_use = 0
        self.stats = ServerStats()

    def submit(self, r: Request) -> None:
        self.waiting.append(r)

    def _admit(self, now: float) -> bool:
        changed = False
        keep = []
        for r in self.waiting:
            need = device_bytes_estimate(r)
            if len(self.active) < self.max_members and self.bytes_in_use + need <= self.budget:
                case = r.case()
                r.domain = RK3(self.ctx, case.grid, r.params(case), case)
                r.admitted = now
                self.bytes_in_use += r.domain.device_bytes
                self.active.append(r)
                self.stats.admitted += 1
                changed = True
            else:
                keep.append(r)
        self.waiting = keep
        self.stats.peak_bytes = max(self.stats.peak_bytes, self.bytes_in_use)
        self.stats.peak_members = max(self.stats.peak_members, len(self.active))
        return changed

    def _retire(self, now: float) -> bool:
        done = [r for r in self.active if r.taken >= r.steps]
        if not done:
            return False
        for r in done:
            r.done = now
            self.bytes_in_use -= r.domain.device_bytes
            self.stats.latencies.append(r.done - r.arrival)
            self.stats.completed += 1
        self.active = [r for r in self.active if r.taken < r.steps]
        self.finished.extend(done)
        return True

    def tick(self, now: float) -> None:
        """one step boundary: retire, admit, re-form the batch, advance every active forecast one step"""
        changed = self._retire(now)
        changed = self._admit(now) or changed
        if not self.active:
            self.batch = None
            return
        if changed or self.batch is None:
            if self.batch is not None:
                self.batch.close()
            self.batch = Batch(self.ctx, [r.domain for r in self.active])
        self.batch.step(1)
        self.batch.sync()  # host and device in step: the next boundary's decisions see the device's progress
        for r in self.active:
            r.taken += 1
            self.stats.point_steps += r.points
        self.stats.ticks += 1

    def release_finished(self, keep=frozenset()) -> None:
        """free the device state of retired forecasts; the final u, v, w, t, mu of those in `keep` (request ids)
        are read back first into r.result"""
        for r in self.finished:
            if r.domain is not None:
                if r.rid in keep:
                    r.result = {fid: r.domain.field(fid) for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_MU)}
                r.domain.close()
                r.domain = None


def poisson_trace(seconds: float, rate: float, seed: int, scale: float = 1.0, precision: int = A.MI_FP64):
    rng = np.random.default_rng(seed)
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        nx, ny = (int(round(rng.uniform(50, 300) * scale)) for _ in range(2))
        nz = int(rng.choice([35, 44, 50]))
        dx = float(rng.uniform(3000.0, 15000.0))
        hours = rng.uniform(1.0, 6.0)
        dt = 6.0 * dx / 1000.0
        steps = max(1, int(round(hours * 3600.0 / dt * scale)))
        out.append(Request(rid, max(nx, 8), max(ny, 8), nz, dx, steps, seed=int(rng.integers(1 << 31)), arrival=t,
                           precision=precision))
        rid += 1
    return out


def run_trace(reqs: list[Request], hbm_budget: int, ctx: Context | None = None, keep=frozenset()) -> dict:
    """replay the trace in real time (arrivals by the wall clock) until every request has finished; the final
    states of the requests in `keep` stay in their Request.result"""
    ctx = ctx or Context(0)
    t0 = time.perf_counter()
    srv = ForecastServer(ctx, hbm_budget, clock=lambda: time.perf_counter() - t0)
    pending = sorted(reqs, key=lambda r: r.arrival)
    while pending or srv.waiting or any(r.taken < r.steps for r in srv.active):
        now = time.perf_counter() - t0
        while pending and pending[0].arrival <= now:
            srv.submit(pending.pop(0))
        if not srv.waiting and not srv.active:
            time.sleep(max(0.0, min(0.05, pending[0].arrival - now)))
            continue
        srv.tick(now)
        srv.release_finished(keep)
    srv.tick(time.perf_counter() - t0)  # the last boundary retires what the last step finished
    srv.release_finished(keep)
    wall = time.perf_counter() - t0
    lat = np.array(srv.stats.latencies)
    return {
        "requests": len(reqs),
        "completed": srv.stats.completed,
        "p50_time_to_forecast_s": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_time_to_forecast_s": float(np.percentile(lat, 99)) if lat.size else None,
        "mean_time_to_forecast_s": float(lat.mean()) if lat.size else None,
        "throughput_grid_point_steps_per_s": srv.stats.point_steps / wall,
        "wall_s": wall,
        "ticks": srv.stats.ticks,
        "peak_members": srv.stats.peak_members,
        "peak_device_bytes": srv.stats.peak_bytes,
        "hbm_budget_bytes": hbm_budget,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--seconds", type=float, default=30.0, help="length of the arrival trace")
    ap.add_argument("--rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--scale", type=float, default=1.0, help="scale of the domain sizes and forecast lengths")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32")
    ap.add_argument("--budget-gib", type=float, default=200.0, help="HBM budget for resident forecasts")
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    prec = A.MI_FP32_BULK if a.precision == "fp32" else A.MI_FP64
    reqs = poisson_trace(a.seconds, a.rate, a.seed, a.scale, prec)
    res = run_trace(reqs, int(a.budget_gib * (1 << 30)))
    res.update({"trace_seconds": a.seconds, "rate_per_s": a.rate, "scale": a.scale, "precision": a.precision})
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
