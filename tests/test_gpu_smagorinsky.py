"""km_opt 4 (the 2-D Smagorinsky horizontal K with diff_opt 1; the reference's DiffusionDriver::prepare,
src/dyn/diffusion_driver.cpp:92-138, and KSmag2dKm, src/dyn/diffusion/km_kh.hpp:47-102): the HIP path against the
oracle, bit for bit in FP64 -- the eddy viscosities of each step, the held tendencies, the state after each phase
and over graph-replayed steps.  The oracle's km_opt 4 is pinned by the closed forms of
test_oracle_smagorinsky.py.

Cases: the reference's straka density current (x-z, periodic); a 3-D bubble with moist species and a tracer,
periodic; the same with open boundaries in x and y and with symmetric ones; the real-data-shaped hills with
specified boundaries, terrain, map factors, both diffusions (second order with km_opt 4 and sixth order) and moist
species.  FP32-bulk stays within 1e-3 relative L2 of FP64 on the hills case."""
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
TENDF = (A.MI_RU_TENDF, A.MI_RV_TENDF, A.MI_RW_TENDF, A.MI_T_TENDF)
KS = (A.MI_XKMH, A.MI_XKHH)
SMAG = dict(diff_opt=1, km_opt=4, khdif=0.0, kvdif=10.0)


def _cases():
    return {
        "straka": (lambda: Case.straka(128, 32, dx=200.0, dt=1.0), dict(SMAG, pr_inv=1.0, kvdif=75.0), False),
        "bubble_moist": (lambda: Case.bubble3d(22, 18, 16, dx=1000.0, ztop=8000.0, dt=6.0, seed=5, bubble_zc=3000.0),
                         dict(SMAG, time_step_sound=6, n_moist=2, n_scalars=3, c_s=0.3), False),
        "bubble_open": (lambda: Case.bubble3d(24, 20, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=6, bubble_zc=3000.0,
                                              bc=A.MI_BC_OPEN),
                        dict(SMAG, time_step_sound=6, n_moist=1, n_scalars=2), False),
        "bubble_symmetric": (lambda: Case.bubble3d(20, 16, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=7,
                                                   bubble_zc=3000.0, bc=A.MI_BC_SYMMETRIC),
                             dict(SMAG, time_step_sound=6, pr_inv=2.0), False),
        "hills_spec": (lambda: Case.hills(24, 20, 16, dx=15000.0, dt=90.0, seed=4, bc=A.MI_BC_SPECIFIED),
                       dict(SMAG, time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=3, kvdif=5.0, diff_6th_opt=1),
                       True),
    }


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _pair(ctx, name, precision=A.MI_FP64):
    make, kw, spec = _cases()[name]
    case = make()
    prm = case.params(precision=precision, **kw)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm) if precision == A.MI_FP64 else None
    g = case.grid
    if spec:
        bdy = drifting_boundary(case, seed=13, q_species=(0,))
        gpu.load_boundary(bdy)
        if orc:
            orc.load_boundary(bdy)
    # a sheared flow so the deformation (and K) varies: u, v from smooth waves added to the case's state
    z, y, x = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    for fid, amp in ((A.MI_U, 3.0), (A.MI_V, -2.0)):
        f = gpu.field(fid)
        zz, yy, xx = np.indices(f.shape)
        f = f + amp * np.sin(2 * np.pi * xx / g.nx + 0.3 * zz) * np.cos(2 * np.pi * yy / max(g.ny, 1) + 0.2)
        if g.ny == 1 and fid == A.MI_V:
            continue
        gpu.set_field(fid, f)
        if orc:
            orc.set_field(fid, f)
    for n in range(1, prm.n_scalars):
        tr = 1e-3 * (1.0 + np.cos(0.4 * x + 0.3 * y + n) * np.exp(-0.1 * z))
        gpu.set_field(A.MI_Q + n, tr)
        if orc:
            orc.set_field(A.MI_Q + n, tr)
    if prm.n_moist > 0 and not case.moist:
        qv = 0.012 * np.exp(-z / 12.0) * (1.0 + 0.1 * np.sin(0.5 * x + 0.2 * y))
        gpu.set_field(A.MI_Q, qv)
        if orc:
            orc.set_field(A.MI_Q, qv)
    return case, prm, gpu, orc


def _compare(gpu, orc, fids, where):
    for fid in fids:
        a, b = gpu.field(fid), orc.field(fid)
        assert bitwise_equal(a, b), f"{where}: {A.FIELD_NAMES.get(fid, fid)}: {first_diff(a, b)}"


@pytest.mark.parametrize("name", list(_cases()))
def test_smagorinsky_phases_bitwise(ctx, name):
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[smagorinsky] phases {name} grid {case.grid}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    tendf = TENDF if case.grid.ny > 1 else (A.MI_RU_TENDF, A.MI_RW_TENDF, A.MI_T_TENDF)
    for step in range(2):
        for d in (gpu, orc):
            d.step_begin()
        for rk in (1, 2, 3):
            for d in (gpu, orc):
                d.rk_prep(rk)
                d.rk_tendency(rk)
            if rk == 1:
                _compare(gpu, orc, KS, f"step {step} eddy viscosities")
                _compare(gpu, orc, tendf, f"step {step} held tendencies")
            for d in (gpu, orc):
                d.small_steps(rk)
            _compare(gpu, orc, STATE, f"step {step} rk {rk} acoustic steps")
            for d in (gpu, orc):
                d.scalar_transport(rk)
            _compare(gpu, orc, qs, f"step {step} rk {rk} scalars")
            for d in (gpu, orc):
                d.calc_p_rho_phi(rk)
        for d in (gpu, orc):
            d.step_end()
        _compare(gpu, orc, STATE + qs, f"step {step} end")
    kmh = gpu.field(A.MI_XKMH)
    assert kmh.max() > 0 and kmh.std() > 0  # K set and varying


@pytest.mark.parametrize("name", ["straka", "hills_spec"])
def test_smagorinsky_graph_steps_bitwise(ctx, name):
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[smagorinsky] graph steps {name}", flush=True)
    for chunk in (1, 4, 5):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE + KS, f"after a chunk of {chunk} steps")


def test_smagorinsky_fp32_close_to_fp64(ctx):
    print("[smagorinsky] hills_spec FP32-bulk vs FP64, 20 steps", flush=True)
    _, _, g64, _ = _pair(ctx, "hills_spec", A.MI_FP64)
    _, _, g32, _ = _pair(ctx, "hills_spec", A.MI_FP32_BULK)
    g64.step(20)
    g32.step(20)
    for name, v in W.error_report(g32, g64, True).items():
        assert v["rel_l2"] <= 1e-3, (name, v)


def test_smagorinsky_options_refused(ctx):
    case = Case.bubble3d(16, 12, 10, dx=1000.0, ztop=8000.0, dt=6.0, seed=1)
    for kw, msg in ((dict(diff_opt=1, km_opt=2), "km_opt"), (dict(diff_opt=1, km_opt=3), "km_opt"),
                    (dict(diff_opt=1, km_opt=4, c_s=-0.1), "c_s")):
        with pytest.raises(Exception, match=msg):
            RK3(ctx, case.grid, case.params(**kw), case)
