"""Continuous-batching forecast server on one GPU.

Requests arrive over time, each one forecast: a domain (its own size, resolution and options), an initial
state and a length in steps.  The server keeps every admitted forecast's state resident on the device and
advances all of them together, one step each per tick: each forecast replays its own captured step graph on
one of a few streams, so the members' kernels overlap and a change of membership captures nothing.  The host
does not wait for the device at a tick (events recorded after each tick bound how far it runs ahead and time
each forecast's completion), and domains are built on background threads while the GPU steps the others.  At
each step boundary it retires the forecasts that are done, admits waiting ones while the HBM budget allows (a
domain's device bytes predicted from those already built), and re-forms the batch.  It reports time-to-forecast
(arrival to completion) percentiles, the aggregate throughput in grid-point-steps/s, the offered load and the
queue length.

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
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from .dyn import Case, Context, RK3


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
    n_scalars: int = 0
    slot: int = 0            # the server's stream the forecast runs on
    reserved: int = 0        # device bytes counted against the budget
    last_event: object = None
    case_s: float = 0.0      # host time of building the case (initial state) and of creating and loading the domain
    domain_s: float = 0.0

    @property
    def points(self) -> int:
        return self.nx * self.ny * self.nz

    def case(self):
        dt = 6.0 * self.dx / 1000.0
        return Case.bubble3d(self.nx, self.ny, self.nz, dx=self.dx, ztop=20000.0, dt=dt, seed=self.seed,
                             noise_amp=0.05, bubble_zc=2000.0)

    def params(self, case):
        return case.params(time_step_sound=self.n_sound, damp_opt=3, zdamp=5000.0, dampcoef=0.2, precision=self.precision)


def alloc_elems(nx: int, ny: int, nz: int) -> int:
    """elements of one 3-D allocation of the library's layout: rows of a pitch that is a multiple of 32 elements
    (8 of padding and 3 of halo west of i = 0, the U points and 3 east), nz + 1 levels, ny + 1 + 6 rows"""
    pitch = (8 + nx + 1 + 3 + 31) // 32 * 32
    rows = (ny + 1 + 6) if ny > 1 else 1
    return pitch * (nz + 1) * rows


class MemoryModel:
    """Device bytes of a domain before it exists: bytes per allocation element of the domains already created with
    the same precision and scalar count (a domain's memory is a fixed number of whole allocations), and before the
    first one a bound from above (every array counted as FP64)."""

    def __init__(self):
        self.ratio: dict = {}

    def estimate(self, r: "Request") -> int:
        n = alloc_elems(r.nx, r.ny, r.nz)
        key = (r.precision, r.n_scalars)
        return int(n * self.ratio[key] * 1.01) if key in self.ratio else int(n * 8 * 80)

    def learn(self, r: "Request", device_bytes: int) -> None:
        self.ratio[(r.precision, r.n_scalars)] = device_bytes / alloc_elems(r.nx, r.ny, r.nz)


def device_bytes_estimate(r: "Request") -> int:
    """the bound from above of MemoryModel (no domain of the kind created yet)"""
    return MemoryModel().estimate(r)


@dataclass
class ServerStats:
    admitted: int = 0
    completed: int = 0
    ticks: int = 0
    point_steps: float = 0.0
    peak_bytes: int = 0
    peak_members: int = 0
    latencies: list = field(default_factory=list)
    queue_samples: list = field(default_factory=list)   # (waiting + preparing) at each tick
    active_samples: list = field(default_factory=list)  # active forecasts at each tick


class ForecastServer:
    """Continuous batching on one GPU.  Admitted forecasts are spread over `streams` contexts, each on a stream of
    its own (default priority), so the members' steps overlap on the device; a tick enqueues one step of every
    active forecast and returns without waiting for the device (at most `max_ahead` ticks are in flight: the host
    then waits for the oldest).  Domains are built (case generation, allocation, upload) by `prepare_workers`
    background threads while the GPU steps the others (0: built on the tick path).  A forecast's completion time is
    when the device finishes its last step (an event recorded on its stream after that step)."""

    def __init__(self, ctx: Context | None, hbm_budget: int, max_members: int = 64, clock=time.perf_counter,
                 streams: int = 4, prepare_workers: int = 0, max_ahead: int = 2):
        import torch
        self._torch = torch
        self.clock = clock
        self.budget = hbm_budget
        self.max_members = max_members
        self.max_ahead = max_ahead
        dev = torch.cuda.current_device()
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(max(1, streams))]
        self.ctxs = [Context(dev, s.cuda_stream) for s in self.streams]
        self.members_on = [0] * len(self.ctxs)
        self.pool = ThreadPoolExecutor(prepare_workers) if prepare_workers > 0 else None
        self.mem = MemoryModel()
        self.waiting: list[Request] = []
        self.preparing: list[tuple[Request, object]] = []
        self.ready: list[Request] = []
        self.active: list[Request] = []
        self.draining: list[Request] = []
        self.finished: list[Request] = []
        self.inflight: list[list] = []  # per tick: the events of the streams it used
        self.bytes_in_use = 0
        self.stats = ServerStats()

    def submit(self, r: Request) -> None:
        self.waiting.append(r)

    def _prepare(self, r: Request) -> Request:
        t0 = time.perf_counter()
        case = r.case()
        t1 = time.perf_counter()
        r.domain = RK3(self.ctxs[r.slot], case.grid, r.params(case), case)
        r.case_s, r.domain_s = t1 - t0, time.perf_counter() - t1
        return r

    def _start_preparing(self, now: float) -> None:
        keep = []
        for r in self.waiting:
            need = self.mem.estimate(r)
            n_live = len(self.active) + len(self.preparing) + len(self.ready)
            if n_live < self.max_members and self.bytes_in_use + need <= self.budget:
                r.reserved = need
                self.bytes_in_use += need
                r.slot = min(range(len(self.ctxs)), key=lambda q: self.members_on[q])
                self.members_on[r.slot] += 1
                if self.pool is None:
                    self.ready.append(self._prepare(r))
                else:
                    self.preparing.append((r, self.pool.submit(self._prepare, r)))
            else:
                keep.append(r)
        self.waiting = keep

    def _collect_prepared(self) -> None:
        still = []
        for r, fut in self.preparing:
            if fut.done():
                fut.result()  # raises what the preparation raised
                self.ready.append(r)
            else:
                still.append((r, fut))
        self.preparing = still

    def _admit(self, now: float) -> bool:
        changed = False
        for r in self.ready:
            real = r.domain.device_bytes
            self.mem.learn(r, real)
            self.bytes_in_use += real - r.reserved
            r.reserved = real
            r.admitted = now
            self.active.append(r)
            self.stats.admitted += 1
            changed = True
        self.ready = []
        self.stats.peak_bytes = max(self.stats.peak_bytes, self.bytes_in_use)
        self.stats.peak_members = max(self.stats.peak_members, len(self.active))
        return changed

    def _retire(self, now: float) -> bool:
        """forecasts whose last step is enqueued leave the batch; they are done when their stream passes the
        event recorded after that step"""
        moved = [r for r in self.active if r.taken >= r.steps]
        if moved:
            self.active = [r for r in self.active if r.taken < r.steps]
            self.draining.extend(moved)
        still = []
        for r in self.draining:
            if r.last_event is None or r.last_event.query():
                r.done = self.clock()
                self.stats.latencies.append(r.done - r.arrival)
                self.stats.completed += 1
                self.bytes_in_use -= r.reserved
                self.members_on[r.slot] -= 1
                self.finished.append(r)
            else:
                still.append(r)
        self.draining = still
        return bool(moved)

    def tick(self, now: float) -> None:
        """one step boundary: collect built domains, retire, admit, enqueue one step of every active forecast"""
        self._collect_prepared()
        self._retire(now)
        self._start_preparing(now)
        self._admit(now)
        self.stats.queue_samples.append(len(self.waiting) + len(self.preparing))
        self.stats.active_samples.append(len(self.active))
        if not self.active:
            return
        used = set()
        for r in self.active:
            r.domain.step(1)
            r.taken += 1
            used.add(r.slot)
            self.stats.point_steps += r.points
        events = {}
        for q in used:
            ev = self._torch.cuda.Event()
            ev.record(self.streams[q])
            events[q] = ev
        for r in self.active:
            if r.taken >= r.steps:
                r.last_event = events[r.slot]
        self.inflight.append(list(events.values()))
        while len(self.inflight) > self.max_ahead:  # bound the run-ahead of the host
            for ev in self.inflight.pop(0):
                ev.synchronize()
        self.stats.ticks += 1

    def drain(self) -> None:
        """wait for every enqueued step and retire what finished"""
        for evs in self.inflight:
            for ev in evs:
                ev.synchronize()
        self.inflight = []
        self._retire(self.clock())

    def release_finished(self, keep=frozenset()) -> None:
        """free the device state of finished forecasts; the final u, v, w, t, mu of those in `keep` (request ids)
        are read back first into r.result"""
        for r in self.finished:
            if r.domain is not None:
                if r.rid in keep:
                    r.result = {fid: r.domain.field(fid) for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_MU)}
                r.domain.close()
                r.domain = None

    def close(self) -> None:
        if self.pool is not None:
            self.pool.shutdown(wait=True)
        self.drain()


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


def run_trace(reqs: list[Request], hbm_budget: int, ctx: Context | None = None, keep=frozenset(), streams: int = 4,
              prepare_workers: int = 2, progress=None, max_wall: float = 0.0, max_ahead: int = 2) -> dict:
    """replay the trace in real time (arrivals by the wall clock) until every request has finished, or until
    `max_wall` seconds (> 0) have passed: then no further step is enqueued, and the requests not finished are
    counted as unfinished (the latencies are those of the finished ones); the final states of the requests in
    `keep` stay in their Request.result"""
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0  # noqa: E731
    srv = ForecastServer(ctx, hbm_budget, clock=clock, streams=streams, prepare_workers=prepare_workers,
                         max_ahead=max_ahead)
    pending = sorted(reqs, key=lambda r: r.arrival)
    last_report = 0.0
    busy_ticks = 0
    stopped = False
    while pending or srv.waiting or srv.preparing or srv.ready or srv.active or srv.draining:
        now = clock()
        if max_wall > 0 and now > max_wall:
            stopped = True
            break
        while pending and pending[0].arrival <= now:
            srv.submit(pending.pop(0))
        if not (srv.waiting or srv.preparing or srv.ready or srv.active or srv.draining):
            time.sleep(max(0.0, min(0.05, pending[0].arrival - now)))
            continue
        if not srv.active and not srv.ready:
            srv.drain()
            srv._collect_prepared()
            if not srv.ready and srv.preparing:
                time.sleep(0.002)
        srv.tick(now)
        busy_ticks += 1 if srv.active else 0
        srv.release_finished(keep)
        if progress and now - last_report > 30.0:
            last_report = now
            progress(f"serving: t={now:.0f}s completed {srv.stats.completed}/{len(reqs)} active {len(srv.active)} "
                     f"waiting {len(srv.waiting) + len(srv.preparing)}")
    srv.close()
    srv.release_finished(keep)
    wall = clock()
    if stopped:  # the device state of the forecasts left unfinished (their enqueued steps have completed)
        for r in reqs:
            if r.domain is not None:
                r.domain.close()
                r.domain = None
    lat = np.array(srv.stats.latencies)
    q = np.array(srv.stats.queue_samples) if srv.stats.queue_samples else np.zeros(1)
    act = np.array(srv.stats.active_samples) if srv.stats.active_samples else np.zeros(1)
    offered = sum(r.points * r.steps for r in reqs)
    trace_s = max((r.arrival for r in reqs), default=0.0)
    return {
        "requests": len(reqs),
        "completed": srv.stats.completed,
        "p50_time_to_forecast_s": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_time_to_forecast_s": float(np.percentile(lat, 99)) if lat.size else None,
        "mean_time_to_forecast_s": float(lat.mean()) if lat.size else None,
        "throughput_grid_point_steps_per_s": srv.stats.point_steps / wall,
        "offered_grid_point_steps_per_s": offered / trace_s if trace_s > 0 else None,
        "wall_s": wall,
        "trace_span_s": trace_s,
        "ticks": srv.stats.ticks,
        "queue_mean": float(q.mean()),
        "queue_max": int(q.max()),
        "queue_last_quarter_mean": float(q[3 * len(q) // 4:].mean()),
        "active_mean": float(act.mean()),
        "peak_members": srv.stats.peak_members,
        "peak_device_bytes": srv.stats.peak_bytes,
        "hbm_budget_bytes": hbm_budget,
        "streams": streams,
        "drain_s": wall - trace_s,
        # stable: the backlog left when arrivals stop clears within a tenth of the trace's span (at least 10 s);
        # a queue that grows while requests arrive (offered work above what the GPU serves) leaves a longer one
        "queue_stable": bool(not stopped and lat.size and wall - trace_s <= max(0.1 * trace_s, 10.0)),
        "stopped_at_max_wall": stopped,
        "unfinished": len(reqs) - srv.stats.completed,
        # host time per request of building its case and of creating and loading its domain (background threads)
        "case_s_mean": float(np.mean([r.case_s for r in reqs])) if reqs else None,
        "domain_s_mean": float(np.mean([r.domain_s for r in reqs])) if reqs else None,
        "domain_s_max": float(max(r.domain_s for r in reqs)) if reqs else None,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--seconds", type=float, default=30.0, help="length of the arrival trace")
    ap.add_argument("--rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--scale", type=float, default=1.0, help="scale of the domain sizes and forecast lengths")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32")
    ap.add_argument("--budget-gib", type=float, default=200.0, help="HBM budget for resident forecasts")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--progress", action="store_true", help="report the replay's progress on stderr every 30 s")
    ap.add_argument("--streams", type=int, default=4, help="streams the forecasts are spread over")
    ap.add_argument("--max-ahead", type=int, default=2, help="ticks the host may enqueue ahead of the device")
    ap.add_argument("--max-wall", type=float, default=0.0,
                    help="stop enqueuing steps after this many seconds (0: run until every request has finished)")
    a = ap.parse_args(argv)
    prec = A.MI_FP32_BULK if a.precision == "fp32" else A.MI_FP64
    reqs = poisson_trace(a.seconds, a.rate, a.seed, a.scale, prec)
    report = (lambda m: print(m, file=sys.stderr, flush=True)) if a.progress else None
    res = run_trace(reqs, int(a.budget_gib * (1 << 30)), progress=report, max_wall=a.max_wall, streams=a.streams,
                    max_ahead=a.max_ahead)
    res["max_ahead"] = a.max_ahead
    res.update({"trace_seconds": a.seconds, "rate_per_s": a.rate, "scale": a.scale, "precision": a.precision})
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
