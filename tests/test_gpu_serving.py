"""The continuous-batching server: forecasts admitted at different step boundaries, batched with others of
other sizes and precisions, give the bits they give alone; latency and throughput are reported."""
import time

import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Context, RK3
from cppwrf_amd.serving import ForecastServer, Request, poisson_trace, run_trace

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def test_results_same_alone_and_in_the_mix(ctx):
    print("[serving] alone vs in the mix", flush=True)
    reqs = [Request(0, 24, 20, 16, 3000.0, 9, seed=1), Request(1, 40, 18, 20, 5000.0, 5, seed=2),
            Request(2, 30, 30, 16, 4000.0, 12, seed=3, precision=A.MI_FP32_BULK), Request(3, 20, 24, 16, 3000.0, 4, seed=4)]
    srv = ForecastServer(ctx, hbm_budget=1 << 34, clock=lambda: 0.0)
    srv.submit(reqs[0])
    srv.submit(reqs[1])
    srv.tick(0.0)
    srv.tick(1.0)
    srv.submit(reqs[2])  # joins at a later step boundary
    srv.tick(2.0)
    srv.submit(reqs[3])
    for t in range(3, 30):
        srv.tick(float(t))
    assert all(r.taken == r.steps for r in reqs)
    for r in reqs:
        case = r.case()
        alone = RK3(ctx, case.grid, r.params(case), case)
        alone.step(r.steps)
        for fid in STATE:
            a, m = alone.field(fid), r.domain.field(fid)
            assert np.array_equal(a.view(np.uint64), m.view(np.uint64)), (r.rid, A.FIELD_NAMES[fid])


def test_background_builds_beside_captures(ctx):
    """domains built by background threads on the streams on which the tick captures other domains' first steps
    give the bits they give alone (a capture on a shared stream would record the builders' fills and uploads)"""
    print("[serving] background builds beside captures", flush=True)
    reqs = [Request(q, 20 + 4 * (q % 4), 18 + 2 * (q % 3), 16, 3000.0 + 500.0 * q, 6 + q % 5, seed=10 + q,
                    precision=A.MI_FP32_BULK if q % 3 == 2 else A.MI_FP64) for q in range(12)]
    srv = ForecastServer(ctx, hbm_budget=1 << 34, clock=lambda: 0.0, streams=2, prepare_workers=2)
    pending = list(reqs)
    t0 = time.perf_counter()
    t = 0
    while pending or srv.waiting or srv.preparing or srv.ready or srv.active or srv.draining:
        assert time.perf_counter() - t0 < 90.0
        if pending and t % 2 == 0:
            srv.submit(pending.pop(0))
        if not srv.active and not srv.ready:
            time.sleep(0.002)
        srv.tick(float(t))
        t += 1
    srv.close()
    assert all(r.taken == r.steps for r in reqs)
    for r in reqs:
        case = r.case()
        alone = RK3(ctx, case.grid, r.params(case), case)
        alone.step(r.steps)
        for fid in STATE:
            a, m = alone.field(fid), r.domain.field(fid)
            assert np.array_equal(a.view(np.uint64), m.view(np.uint64)), (r.rid, A.FIELD_NAMES[fid])
        alone.close()
        r.domain.close()


def test_short_poisson_trace(ctx):
    print("[serving] short Poisson trace", flush=True)
    reqs = poisson_trace(seconds=3.0, rate=2.0, seed=5, scale=0.2, precision=A.MI_FP32_BULK)
    res = run_trace(reqs, hbm_budget=8 << 30, ctx=ctx)
    print(res, flush=True)
    assert res["completed"] == len(reqs)
    assert res["p99_time_to_forecast_s"] >= res["p50_time_to_forecast_s"] > 0


def test_trace_stopped_at_max_wall(ctx):
    print("[serving] a trace stopped at its wall-clock limit", flush=True)
    reqs = poisson_trace(seconds=4.0, rate=3.0, seed=7, scale=0.5, precision=A.MI_FP32_BULK)
    for r in reqs:
        r.steps = 100000  # none can finish within the limit
    res = run_trace(reqs, hbm_budget=8 << 30, ctx=ctx, max_wall=2.0)
    print({k: res[k] for k in ("requests", "completed", "unfinished", "stopped_at_max_wall", "wall_s")}, flush=True)
    assert res["stopped_at_max_wall"] and not res["queue_stable"]
    assert res["completed"] == 0 and res["unfinished"] == len(reqs) > 0
    assert res["wall_s"] < 30.0
    assert all(r.domain is None for r in reqs)  # the unfinished forecasts' device state is released
