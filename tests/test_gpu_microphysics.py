"""Kessler microphysics after the RK loop (the reference's PhysicsCoupling with mp_physics = 1) on the GPU against the
oracle, bit for bit in FP64: the squall2d case through its first rain (condensation by step 100, rain at the ground
by step 200), the split fall of heavy rain at a 60 s step, a 3-D domain with specified boundaries, and a checkpoint
that carries the heating and the precipitation.  The oracle's coupling is pinned to the reference's squall2d golden
(test_oracle_golden.py) and its scheme by test_oracle_microphysics.py."""
import numpy as np
import pytest

from conftest import bitwise_equal, first_diff
from cppwrf_amd import _abi as A
from cppwrf_amd import workloads as W
from cppwrf_amd.boundary import drifting_boundary
from cppwrf_amd.dyn import Case, Context, RK3
from oracle_lib import OracleDomain

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)
MOIST = (A.MI_Q, A.MI_Q + 1, A.MI_Q + 2, A.MI_RAINNC, A.MI_RAINNCV, A.MI_H_DIABATIC)
KW = dict(n_moist=3, n_scalars=3, mp_physics=1)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _compare(gpu, orc, fids, where):
    for fid in fids:
        a, b = gpu.field(fid), orc.field(fid)
        assert bitwise_equal(a, b), f"{where}: {A.FIELD_NAMES.get(fid, fid)}: {first_diff(a, b)}"


def test_squall2d_through_the_first_rain(ctx):
    case = Case.squall2d()
    prm = case.params(**KW)
    print("[microphysics] squall2d 120x1x40, 300 graph-replayed steps against the oracle", flush=True)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    done = 0
    for chunk in (1, 49, 50, 50, 50, 100):
        gpu.step(chunk)
        orc.step(chunk)
        done += chunk
        _compare(gpu, orc, STATE + MOIST, f"after {done} steps")
    assert gpu.field(A.MI_Q + 1).max() > 1e-4 and gpu.field(A.MI_RAINNC).max() > 0.1  # cloud and rain formed


def _after_rk(gpu, orc):
    for d in (gpu, orc):
        d.step_begin()
        for rk in (1, 2, 3):
            d.rk_prep(rk)
            d.rk_tendency(rk)
            d.small_steps(rk)
            d.scalar_transport(rk)
            d.calc_p_rho_phi(rk)


def test_split_fall_of_heavy_rain(ctx):
    """one 60 s step's end with 0.08 kg/kg of rain in the lowest six layers: the fall splits into sub-steps"""
    case = Case.squall2d(dt=60.0)
    prm = case.params(**KW)
    print("[microphysics] heavy rain at dt 60 s", flush=True)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    gpu.step(2)
    orc.step(2)
    _after_rk(gpu, orc)
    qr = orc.field(A.MI_Q + 2)
    qr[:6, 0, 30:40] = 0.08
    qc = orc.field(A.MI_Q + 1)
    qc[4:12, 0, 60:80] = 2.5e-3
    for d in (gpu, orc):
        d.set_field(A.MI_Q + 2, qr)
        d.set_field(A.MI_Q + 1, qc)
        d.step_end()
    _compare(gpu, orc, STATE + MOIST, "the step's end")
    rain = gpu.field(A.MI_RAINNC)[0, 0, 30:40]
    assert np.all(rain > 1.5 * gpu.field(A.MI_RAINNCV)[0, 0, 30:40])
    # (the dynamics is not stable at this step: the comparison stops at the step's end)


def test_specified_3d_with_rain(ctx):
    """a 3-D moist domain with specified boundaries: Kessler skips the spec zone, calc_p_rho_phi after it runs on
    the interior"""
    case = Case.hills(24, 20, 16, dx=15000.0, dt=90.0, seed=6, bc=A.MI_BC_SPECIFIED)
    prm = case.params(time_step_sound=4, damp_opt=3, n_moist=3, n_scalars=3, mp_physics=1)
    print("[microphysics] 3-D hills, specified boundaries", flush=True)
    bdy = drifting_boundary(case, seed=21, q_species=(0,))
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    gpu.load_boundary(bdy)
    orc.load_boundary(bdy)
    gpu.step(1)
    orc.step(1)
    _after_rk(gpu, orc)
    qv = orc.field(A.MI_Q)
    qv[1:6, 4:14, 5:15] *= 1.4  # supersaturate a block
    qr = orc.field(A.MI_Q + 2)
    qr[:4, 6:12, 8:14] = 5e-3
    for d in (gpu, orc):
        d.set_field(A.MI_Q, qv)
        d.set_field(A.MI_Q + 2, qr)
        d.step_end()
    _compare(gpu, orc, STATE + MOIST, "the step's end")
    for chunk in (1, 3):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE + MOIST, f"after {chunk} more steps")
    assert gpu.field(A.MI_RAINNC).max() > 0


def test_checkpoint_carries_heating_and_rain(ctx):
    case = Case.squall2d()
    prm = case.params(**KW)
    print("[microphysics] checkpoint at step 220", flush=True)
    a = RK3(ctx, case.grid, prm, case)
    a.step(220)
    buf = a.save_state()
    a.step(20)
    b = RK3(ctx, case.grid, prm, case)
    b.load_state(buf)
    b.step(20)
    for fid in STATE + MOIST:
        x, y = a.field(fid), b.field(fid)
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), A.FIELD_NAMES.get(fid, fid)


def test_fp32_bulk_close_before_the_rain(ctx):
    case = Case.squall2d()
    print("[microphysics] FP32-bulk vs FP64, 100 steps", flush=True)
    d64 = RK3(ctx, case.grid, case.params(**KW), case)
    d32 = RK3(ctx, case.grid, case.params(precision=A.MI_FP32_BULK, **KW), case)
    d64.step(100)
    d32.step(100)
    rep = W.error_report(d32, d64, True)
    print("[microphysics] FP32-bulk error after 100 steps:", {k: f"{v['rel_l2']:.2e}" for k, v in rep.items()})
    for name, v in rep.items():
        assert v["rel_l2"] <= 1e-3, (name, v)
    d64.step(200)
    d32.step(200)
    r64, r32 = d64.field(A.MI_RAINNC).sum(), d32.field(A.MI_RAINNC).sum()
    print(f"[microphysics] total rain after 300 steps: FP64 {r64:.4f}, FP32-bulk {r32:.4f} mm", flush=True)
    assert abs(r32 - r64) <= 0.05 * r64


def test_microphysics_options_refused(ctx):
    case = Case.squall2d()
    for kw, msg in ((dict(n_moist=3, n_scalars=3, mp_physics=6), "mp_physics"),
                    (dict(n_moist=2, n_scalars=2, mp_physics=1), "qv, qc, qr")):
        with pytest.raises(Exception, match=msg):
            RK3(ctx, case.grid, case.params(**kw), case)
