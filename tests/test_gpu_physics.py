"""The column physics attachment (reference ColumnPhysics, src/dyn/column_physics.hpp:21-34; RK3::step,
rk3.cpp:100-166) on the GPU: a plugin whose part2 sets state-dependent tendencies of u, v, theta and every scalar,
with the advective forcings RTHFTEN / RQVFTEN kept, gives the oracle's trajectory and forcings bit for bit, phase
by phase and through mi_step, with diffusion and with specified boundaries; FP32-bulk stays close to FP64.  The
oracle's side is pinned in test_oracle_physics.py."""
import numpy as np
import pytest

from conftest import bitwise_equal, first_diff
from cppwrf_amd import _abi as A
from cppwrf_amd.boundary import drifting_boundary
from cppwrf_amd.dyn import Case, Context, RK3
from oracle_lib import OracleDomain
from physics_plugin import Plugin

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)
TENDF = (A.MI_RU_TENDF, A.MI_RV_TENDF, A.MI_RW_TENDF, A.MI_T_TENDF)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _cases():
    return {
        "hills_diff": (lambda: Case.hills(22, 20, 16, dx=15000.0, dt=90.0, seed=3),
                       dict(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=3, diff_6th_opt=2), False),
        "hills_plain_nodiff": (lambda: Case.hills(20, 18, 14, dx=15000.0, dt=90.0, seed=5),
                               dict(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=2, moist_adv_opt=0), False),
        "hills_spec": (lambda: Case.hills(24, 20, 16, dx=15000.0, dt=90.0, seed=4, bc=A.MI_BC_SPECIFIED),
                       dict(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=2, diff_6th_opt=1), True),
    }


def _pair(ctx, name, precision=A.MI_FP64):
    make, kw, spec = _cases()[name]
    case = make()
    prm = case.params(precision=precision, **kw)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    if spec:
        bdy = drifting_boundary(case, seed=11, q_species=(0,))
        gpu.load_boundary(bdy)
        orc.load_boundary(bdy)
    pg, po = Plugin(True, True, prm.n_scalars), Plugin(True, True, prm.n_scalars)
    gpu.set_column_physics(pg)
    orc.set_column_physics(po)
    return case, prm, gpu, orc, pg, po


def _compare(gpu, orc, fids, where):
    for fid in fids:
        a, b = gpu.field(fid), orc.field(fid)
        assert bitwise_equal(a, b), f"{where}: {A.FIELD_NAMES[fid]}: {first_diff(a, b)}"


@pytest.mark.parametrize("name", list(_cases()))
def test_physics_phases_bitwise(ctx, name):
    case, prm, gpu, orc, pg, po = _pair(ctx, name)
    print(f"[physics] phases {name}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    for step in range(2):
        gpu.step_begin()
        orc.step_begin()
        for rk in (1, 2, 3):
            gpu.rk_prep(rk)
            orc.rk_prep(rk)
            gpu.rk_tendency(rk)
            orc.rk_tendency(rk)
            if rk == 1:
                _compare(gpu, orc, TENDF, f"step {step} held tendencies")
            _compare(gpu, orc, (A.MI_RU_TEND, A.MI_RV_TEND, A.MI_T_TEND, A.MI_TH_FORCING), f"step {step} rk {rk} tend")
            gpu.small_steps(rk)
            orc.small_steps(rk)
            gpu.scalar_transport(rk)
            orc.scalar_transport(rk)
            _compare(gpu, orc, qs + (A.MI_QV_FORCING,), f"step {step} rk {rk} scalars")
            gpu.calc_p_rho_phi(rk)
            orc.calc_p_rho_phi(rk)
        gpu.step_end()
        orc.step_end()
        _compare(gpu, orc, STATE + qs, f"step {step} end")
    assert pg.calls == po.calls == ["part1", "part2", "after_rk", "after_all"] * 2
    assert pg.tendf_at_part1 == [0.0, 0.0]


@pytest.mark.parametrize("name", ["hills_diff", "hills_spec"])
def test_physics_steps_bitwise(ctx, name):
    case, prm, gpu, orc, pg, po = _pair(ctx, name)
    print(f"[physics] steps {name}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    gpu.step(4)
    orc.step(4)
    _compare(gpu, orc, STATE + qs + (A.MI_TH_FORCING, A.MI_QV_FORCING), "after 4 steps")
    assert pg.calls == po.calls == ["part1", "part2", "after_rk", "after_all"] * 4
    # detached: graph-replayed steps again, the same trajectory as the oracle without the plugin
    gpu.set_column_physics(None)
    orc.set_column_physics(None)
    gpu.step(3)
    orc.step(3)
    _compare(gpu, orc, STATE + qs, "after 3 steps without the plugin")
    assert len(pg.calls) == 16


def test_physics_fp32_close(ctx):
    case, prm, d64, _, _, _ = _pair(ctx, "hills_diff")
    _, _, d32, _, _, _ = _pair(ctx, "hills_diff", precision=A.MI_FP32_BULK)
    print("[physics] fp32", flush=True)
    d64.step(10)
    d32.step(10)
    for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_MU, A.MI_Q, A.MI_TH_FORCING):
        a, b = d32.field(fid), d64.field(fid)
        l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        assert l2 < 1e-3, (A.FIELD_NAMES[fid], l2)


def test_physics_hook_error_is_raised(ctx):
    case, prm, gpu, _, _, _ = _pair(ctx, "hills_plain_nodiff")

    class Bad:
        def part2(self, d):
            raise ValueError("plugin failure")
    gpu.set_column_physics(Bad())
    with pytest.raises(ValueError, match="plugin failure"):
        gpu.step(1)


def test_set_diffusion_between_steps(ctx):
    """RK3::set_diffusion: sixth-order diffusion attached after 2 steps, changed after 4, detached after 6: the
    oracle's trajectory bit for bit, the attached steps as a domain created with the option would take them"""
    case = Case.hills(20, 18, 14, dx=15000.0, dt=90.0, seed=6)
    prm = case.params(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=2)
    gpu, orc = RK3(ctx, case.grid, prm, case), OracleDomain(case, prm)
    print("[physics] set_diffusion", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    for opt, factor in ((0, 0.12), (1, 0.12), (2, 0.25), (0, 0.12)):
        gpu.set_diffusion(opt, factor)
        orc.set_diffusion(opt, factor)
        gpu.step(2)
        orc.step(2)
        _compare(gpu, orc, STATE + qs, f"diffusion {opt} {factor}")
    # attached from the start: the same as a domain created with the option
    fresh = RK3(ctx, case.grid, case.params(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=2, diff_6th_opt=2,
                                           diff_6th_factor=0.25), case)
    late = RK3(ctx, case.grid, prm, case)
    late.set_diffusion(2, 0.25)
    fresh.step(3)
    late.step(3)
    _compare(fresh, late, STATE + qs, "attached at creation or afterwards")
