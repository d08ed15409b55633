"""Specified lateral boundaries (the reference's LateralBC, src/dyn/lateral_bc.cpp) on the GPU: bit for
bit against the oracle phase by phase and over graph-replayed steps, with moisture, terrain, map factors,
diffusion and the boundary clock; FP32-bulk close to FP64.  The oracle's boundary path is checked against
closed forms and the periodic path in test_oracle_specified.py."""
import numpy as np
import pytest

from conftest import bitwise_equal, first_diff
from cppwrf_amd import _abi as A
from cppwrf_amd.boundary import drifting_boundary
from cppwrf_amd.dyn import Case, Context, RK3
from oracle_lib import OracleDomain

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)
PREP = (A.MI_WW, A.MI_ALT, A.MI_PHP)
TEND = (A.MI_RU_TEND, A.MI_RV_TEND, A.MI_RW_TEND, A.MI_T_TEND, A.MI_PH_TEND)
STAGE = STATE + (A.MI_RU_M, A.MI_RV_M, A.MI_WW_M, A.MI_MUTS, A.MI_WW)


def _cases():
    return {
        "bubble": (lambda: Case.bubble3d(20, 16, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=5, bubble_zc=3000.0,
                                         bc=A.MI_BC_SPECIFIED),
                   dict(time_step_sound=6, damp_opt=3, zdamp=3000.0)),
        "hills_moist": (lambda: Case.hills(24, 20, 18, dx=15000.0, dt=90.0, seed=3, bc=A.MI_BC_SPECIFIED),
                        dict(time_step_sound=4, damp_opt=3, n_moist=3, n_scalars=4)),
        "hills_diff": (lambda: Case.hills(22, 21, 16, dx=15000.0, dt=90.0, seed=4, bc=A.MI_BC_SPECIFIED),
                       dict(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=2, diff_6th_opt=2, moist_adv_opt=0)),
    }


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


# the scalars with boundary data: qv in hills_moist (its other species and the tracer take flow_dep_bdy),
# none in hills_diff (qv itself flow-dependent)
Q_DATA = {"bubble": (), "hills_moist": (0,), "hills_diff": ()}


def _pair(ctx, name, precision=A.MI_FP64):
    make, kw = _cases()[name]
    case = make()
    prm = case.params(precision=precision, **kw)
    bdy = drifting_boundary(case, seed=11, q_species=Q_DATA[name])
    gpu = RK3(ctx, case.grid, prm, case)
    gpu.load_boundary(bdy)
    orc = OracleDomain(case, prm)
    orc.load_boundary(bdy)
    g = case.grid
    for n in range(prm.n_moist, prm.n_scalars):  # a passive tracer: a smooth blob, nonzero at the edges
        z, y, x = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
        tr = 1.0 + np.cos(0.4 * x + 0.3 * y) * np.exp(-0.1 * z)
        gpu.set_field(A.MI_Q + n, tr)
        orc.set_field(A.MI_Q + n, tr)
    return case, prm, gpu, orc


def _compare(gpu, orc, fids, where):
    for fid in fids:
        a, b = gpu.field(fid), orc.field(fid)
        assert bitwise_equal(a, b), f"{where}: {A.FIELD_NAMES[fid]}: {first_diff(a, b)}"


@pytest.mark.parametrize("name", list(_cases()))
def test_specified_phases_bitwise(ctx, name):
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[specified] phases {name} grid {case.grid}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    for step in range(2):
        gpu.step_begin()
        orc.step_begin()
        assert gpu.bdy_time == orc.bdy_time == (step + 1) * case.grid.dt
        for rk in (1, 2, 3):
            gpu.rk_prep(rk)
            orc.rk_prep(rk)
            _compare(gpu, orc, PREP, f"step {step} rk {rk} prep")
            gpu.rk_tendency(rk)
            orc.rk_tendency(rk)
            _compare(gpu, orc, TEND, f"step {step} rk {rk} tendencies")
            gpu.small_steps(rk)
            orc.small_steps(rk)
            _compare(gpu, orc, STAGE, f"step {step} rk {rk} acoustic steps")
            gpu.scalar_transport(rk)
            orc.scalar_transport(rk)
            _compare(gpu, orc, qs, f"step {step} rk {rk} scalars")
            gpu.calc_p_rho_phi(rk)
            orc.calc_p_rho_phi(rk)
            _compare(gpu, orc, (A.MI_P, A.MI_AL), f"step {step} rk {rk} calc_p_rho_phi")
        gpu.step_end()
        orc.step_end()
        _compare(gpu, orc, STATE + qs, f"step {step} spec_bdy_final")


@pytest.mark.parametrize("name", ["hills_moist", "hills_diff"])
def test_specified_graph_steps_bitwise(ctx, name):
    """mi_step (graph-replayed) over 12 steps against the oracle; then the boundary interval restarts"""
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[specified] graph steps {name}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    for chunk in (1, 5, 6):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE + qs, f"after a chunk of {chunk} steps")
    assert gpu.bdy_time == pytest.approx(12 * case.grid.dt, abs=0)
    gpu.bdy_time = 0.0
    orc.bdy_time = 0.0
    gpu.step(3)
    orc.step(3)
    _compare(gpu, orc, STATE + qs, "after a new boundary interval")
    assert gpu.check_finite() == -1


def test_specified_autotune_keeps_trajectory(ctx):
    """The autotuner restores the state and the boundary clock: the forecast is the untuned one"""
    case, prm, gpu, orc = _pair(ctx, "bubble")
    print("[specified] autotune", flush=True)
    gpu.step(2)
    orc.step(2)
    gpu.autotune(1)
    assert gpu.bdy_time == 2 * case.grid.dt
    gpu.step(3)
    orc.step(3)
    _compare(gpu, orc, STATE, "after autotune and 3 steps")


def test_specified_rejects_bad_input(ctx):
    case = Case.bubble3d(20, 16, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=5, bc=A.MI_BC_SPECIFIED)
    gpu = RK3(ctx, case.grid, case.params(time_step_sound=6), case)
    with pytest.raises(A.MiError):
        gpu.bdy_set(A.MI_BDY_U, A.MI_BDY_XS, np.zeros(7), np.zeros(7))  # wrong size
    with pytest.raises(A.MiError):
        gpu.bdy_set(A.MI_BDY_Q, A.MI_BDY_XS, np.zeros(16 * 14 * 5), np.zeros(16 * 14 * 5))  # no scalar 0
    periodic = Case.bubble3d(20, 16, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=5)
    gp = RK3(ctx, periodic.grid, periodic.params(time_step_sound=6), periodic)
    with pytest.raises(A.MiError):
        gp.bdy_time = 1.0


def test_specified_fp32_close_to_fp64(ctx):
    """FP32-bulk with specified boundaries stays within the FP32 accuracy bound of FP64 over 40 steps
    (1 simulated hour at dt 90 s): relative L2 of u, v, w, t, ph, mu, qv below 1e-3."""
    _, prm64, g64, _ = _pair(ctx, "hills_moist")
    _, prm32, g32, _ = _pair(ctx, "hills_moist", A.MI_FP32_BULK)
    print("[specified] fp32 vs fp64", flush=True)
    g64.step(40)
    g32.step(40)
    for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_Q):
        a, b = g32.field(fid), g64.field(fid)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 1e-3, (A.FIELD_NAMES[fid], rel)


@pytest.mark.parametrize("bc", [A.MI_BC_SPECIFIED, A.MI_BC_PERIODIC])
def test_calc_p_rho_phi_interior(ctx, bc):
    """SmallStep::calc_p_rho_phi_interior after a change of theta: p and al recomputed away from the spec zone
    (everywhere along a periodic direction), the spec zone's columns keep theirs; bit for bit as the oracle"""
    case = Case.bubble3d(20, 16, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=5, bubble_zc=3000.0, bc=bc)
    prm = case.params(time_step_sound=4)
    gpu, orc = RK3(ctx, case.grid, prm, case), OracleDomain(case, prm)
    print(f"[specified] calc_p_rho_phi_interior bc {bc}", flush=True)
    if bc == A.MI_BC_SPECIFIED:
        bdy = drifting_boundary(case, seed=3)
        gpu.load_boundary(bdy)
        orc.load_boundary(bdy)
    gpu.step(2)
    orc.step(2)
    t = gpu.field(A.MI_T) + 0.5
    p0 = gpu.field(A.MI_P)
    for d in (gpu, orc):
        d.set_field(A.MI_T, t)
        d.calc_p_rho_phi_interior(1)
    _compare(gpu, orc, (A.MI_P, A.MI_AL), "calc_p_rho_phi_interior")
    p1 = gpu.field(A.MI_P)
    if bc == A.MI_BC_SPECIFIED:
        assert np.array_equal(p1[:, :, 0], p0[:, :, 0]) and np.array_equal(p1[:, 0, :], p0[:, 0, :])
    assert not np.array_equal(p1[:, 5:-5, 5:-5], p0[:, 5:-5, 5:-5])
