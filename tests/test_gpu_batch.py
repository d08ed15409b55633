"""Batches: domains of different sizes and options stepped together in one graph per step (mi_batch) give,
member by member, the bits each gives stepped alone; and a member stepped alone in between is picked up."""
import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Batch, Case, Context, RK3

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def members(ctx):
    out = []
    specs = [
        (Case.bubble3d(24, 20, 16, dx=1000.0, ztop=8000.0, dt=6.0, seed=1, bubble_zc=3000.0), dict(time_step_sound=6)),
        (Case.bubble3d(40, 28, 20, dx=2000.0, ztop=10000.0, dt=10.0, seed=2, bubble_zc=3000.0),
         dict(time_step_sound=6, damp_opt=3, zdamp=4000.0)),
        (Case.hills(30, 24, 16, dx=15000.0, dt=90.0, seed=3), dict(time_step_sound=4, n_scalars=2)),
        (Case.bubble3d(24, 20, 16, dx=1000.0, ztop=8000.0, dt=6.0, seed=4, bubble_zc=3000.0),
         dict(time_step_sound=6, precision=A.MI_FP32_BULK, n_moist=1, n_scalars=1)),
    ]
    for case, kw in specs:
        out.append((case, case.params(**kw)))
    return out


def test_batch_equals_alone(ctx):
    print("[batch] 4 members, alone vs batched", flush=True)
    ms = members(ctx)
    alone = [RK3(ctx, c.grid, p, c) for c, p in ms]
    mixed = [RK3(ctx, c.grid, p, c) for c, p in ms]
    for d in alone:
        d.step(7)
    b = Batch(ctx, mixed)
    b.step(3)
    mixed[1].step(1)  # one member on its own in between: its scalar-buffer parity no longer matches
    b.step(3)
    for q in (0, 2, 3):  # every member at 7 steps
        mixed[q].step(1)
    for q, (da, dm) in enumerate(zip(alone, mixed)):
        fids = STATE + ((A.MI_Q,) if ms[q][1].n_scalars else ())
        for fid in fids:
            a, m = da.field(fid), dm.field(fid)
            assert np.array_equal(a.view(np.uint64), m.view(np.uint64)), (q, A.FIELD_NAMES[fid])


def test_batch_graphs_stay_bounded_under_set_dt(ctx):
    """an adaptive time step changes a member's schedule every step: the batch keeps the graphs of the current
    schedules only (one per scalar-buffer parity), not one more per step"""
    print("[batch] set_dt every step, graph count", flush=True)
    ms = members(ctx)[2:4]
    doms = [RK3(ctx, c.grid, p, c) for c, p in ms]
    b = Batch(ctx, doms)
    for s in range(8):
        c, p = ms[0]
        doms[0].set_dt(c.grid.dt * (1.0 - 0.01 * (s % 3)), 4)
        b.step(1)
        assert b.graph_count <= 2, b.graph_count
    b.sync()
    assert all(d.check_finite() == -1 for d in doms)


@pytest.mark.parametrize("precision", [A.MI_FP64, A.MI_FP32_BULK])
def test_fused_batch_equals_alone(ctx, precision):
    """members of one arithmetic and one acoustic step count step as one schedule, each acoustic substep's kernels
    launched once over all members: every member gets the bits it gets alone (periodic, open, specified with moist
    scalars and diffusion, an x-z slice), also after a member stepped alone in between and after set_dt"""
    print(f"[batch] fused schedule, precision {precision}", flush=True)
    specs = [
        (Case.bubble3d(24, 20, 16, dx=1000.0, ztop=8000.0, dt=6.0, seed=1, bubble_zc=3000.0), dict(time_step_sound=4)),
        (Case.bubble3d(40, 28, 20, dx=2000.0, ztop=10000.0, dt=10.0, seed=2, bubble_zc=3000.0, bc=A.MI_BC_OPEN),
         dict(time_step_sound=4, damp_opt=3, zdamp=4000.0)),
        (Case.hills(30, 24, 16, dx=15000.0, dt=90.0, seed=3, bc=A.MI_BC_SPECIFIED),
         dict(time_step_sound=4, n_moist=2, n_scalars=3, diff_6th_opt=1)),
        (Case.bubble3d(70, 1, 20, dx=500.0, ztop=8000.0, dt=3.0, seed=4, bubble_zc=2000.0, bc=A.MI_BC_SYMMETRIC),
         dict(time_step_sound=4)),
    ]
    from cppwrf_amd.boundary import drifting_boundary
    bdy = drifting_boundary(specs[2][0], seed=5)

    def make(case, kw, q):
        d = RK3(ctx, case.grid, case.params(precision=precision, **kw), case)
        if q == 2:
            d.load_boundary(bdy)
        return d
    alone = [make(c, kw, q) for q, (c, kw) in enumerate(specs)]
    mixed = [make(c, kw, q) for q, (c, kw) in enumerate(specs)]
    for d in alone:
        d.step(5)
    b = Batch(ctx, mixed, schedule=A.MI_BATCH_FUSED)
    b.step(2)
    mixed[2].step(1)  # a member alone in between: its scalar buffers' parity changes
    b.step(2)
    for q in (0, 1, 3):
        mixed[q].step(1)
    for q, (da, dm) in enumerate(zip(alone, mixed)):
        fids = STATE + ((A.MI_Q, A.MI_Q + 2) if specs[q][1].get("n_scalars") else ())
        for fid in fids:
            x, y = da.field(fid), dm.field(fid)
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), (q, A.FIELD_NAMES[fid])
    # an adaptive time step on one member: the schedule's tables are rebuilt
    c0 = specs[0][0]
    alone[0].set_dt(c0.grid.dt * 0.9, 4)
    mixed[0].set_dt(c0.grid.dt * 0.9, 4)
    alone[0].step(2)
    for q in (1, 2, 3):
        alone[q].step(2)
    b.step(2)
    for q, (da, dm) in enumerate(zip(alone, mixed)):
        for fid in STATE:
            x, y = da.field(fid), dm.field(fid)
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), ("after set_dt", q, A.FIELD_NAMES[fid])
    assert b.graph_count <= 2
    # members of different acoustic step counts cannot share the fused launches: refused, the batch unchanged
    alone[1].set_dt(specs[1][0].grid.dt, 6)
    b2 = Batch(ctx, [alone[0], alone[1]])
    with pytest.raises(Exception, match="fused schedule"):
        b2.set_schedule(A.MI_BATCH_FUSED)
    b2.step(1)
    b2.close()
    b.close()
