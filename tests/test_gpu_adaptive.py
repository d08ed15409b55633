"""The adaptive-time-step surface of the reference's RK3 (src/dyn/rk3.hpp:40-64, rk3.cpp:74-119): set_dt with
the boundary relaxation weights and the diffusion's dt, and the stage hooks on_cfl / on_first_rk_prep, bit for
bit against the oracle on the GPU."""
import pytest

from conftest import bitwise_equal, first_diff
from cppwrf_amd import _abi as A
from cppwrf_amd.boundary import drifting_boundary
from cppwrf_amd.dyn import Case, Context, RK3
from oracle_lib import OracleDomain

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _compare(a_dom, b_dom, fids, where):
    for fid in fids:
        a, b = a_dom.field(fid), b_dom.field(fid)
        assert bitwise_equal(a, b), f"{where}: {A.FIELD_NAMES[fid]}: {first_diff(a, b)}"


def test_set_dt_bitwise(ctx):
    """Two steps at 90 s with 4 acoustic steps, then three at 60 s with 6: every term that follows dt (the acoustic
    step, w_damping's Courant number, the sixth-order coefficient, the relaxation weights, the boundary clock)
    moves with it, in the oracle's order"""
    case = Case.hills(22, 21, 16, dx=15000.0, dt=90.0, seed=4, bc=A.MI_BC_SPECIFIED)
    prm = case.params(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=2, diff_6th_opt=1, w_damping=1,
                      moist_adv_opt=1)
    bdy = drifting_boundary(case, seed=11, q_species=(0,))
    gpu = RK3(ctx, case.grid, prm, case)
    gpu.load_boundary(bdy)
    orc = OracleDomain(case, prm)
    orc.load_boundary(bdy)
    print("[adaptive] set_dt hills 22x21x16", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    gpu.step(2)
    orc.step(2)
    gpu.set_dt(60.0, 6)
    orc.set_dt(60.0, 6)
    assert gpu.n_sound == 6
    gpu.step(3)
    orc.step(3)
    _compare(gpu, orc, STATE + qs, "3 steps after set_dt")
    assert gpu.bdy_time == orc.bdy_time == 2 * 90.0 + 3 * 60.0
    with pytest.raises(A.MiError):
        gpu.set_dt(60.0, 5)  # odd
    with pytest.raises(A.MiError):
        gpu.set_dt(0.0, 4)


def test_stage_hooks(ctx):
    """on_cfl receives LargeStep::cfl_max after each stage's tendencies (the oracle's maxima at the same points,
    exactly); on_first_rk_prep runs once per step before them; the hooked steps equal the graph-replayed ones,
    and removing the hooks returns to the graph path"""
    case = Case.hills(30, 24, 20, dx=15000.0, dt=90.0, seed=6)
    prm = case.params(time_step_sound=4)
    gpu = RK3(ctx, case.grid, prm, case)
    ref = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    print("[adaptive] stage hooks hills 30x24x20", flush=True)
    calls, preps = [], []
    gpu.on_cfl = lambda v, h: calls.append((v, h))
    gpu.on_first_rk_prep = lambda: preps.append(len(calls))
    gpu.step(2)
    ref.step(2)
    expect = []
    for _ in range(2):
        orc.step_begin()
        for rk in (1, 2, 3):
            orc.rk_prep(rk)
            orc.rk_tendency(rk)
            expect.append(orc.cfl_max())
            orc.small_steps(rk)
            orc.scalar_transport(rk)
            orc.calc_p_rho_phi(rk)
        orc.step_end()
    assert calls == expect and preps == [0, 3], (calls, expect, preps)
    assert all(v > 0 and h > 0 for v, h in calls)
    _compare(gpu, ref, STATE, "hooked steps against graph steps")
    _compare(gpu, orc, STATE, "hooked steps against the oracle")
    gpu.on_cfl = None
    gpu.on_first_rk_prep = None
    gpu.step(1)
    ref.step(1)
    assert len(calls) == 6
    _compare(gpu, ref, STATE, "after the hooks are removed")
