"""The damping layers of the diffusion driver on the GPU against the oracle, bit for bit in FP64: damp_opt 2 (the
Rayleigh relaxation of u, v, w and theta towards the base profiles in the top zdamp metres, the reference's
rk_rayleigh_damp, src/dyn/diffusion/rayleigh_damp.hpp) and damp_opt 1 (the eddy viscosities raised to a damping
profile, cal_dampkm, src/dyn/diffusion/km_kh.hpp:485-568) with km_opt 1 and 4 -- the held tendencies and eddy
viscosities of each step, the state after each phase and over graph-replayed steps.  The oracle's damping is pinned
by the closed forms of test_oracle_damping.py.

Cases: a 3-D bubble in a mean wind, periodic and with symmetric walls; the reference's straka slice (x-z); the
real-data-shaped hills with specified boundaries, terrain, map factors and moisture (damp_opt 2 with the
sixth-order diffusion; damp_opt 1 with the Smagorinsky K); a bubble with dx != dy (damp_opt 1, constant K) and one
with open boundaries (damp_opt 1 raising the edge rows the Smagorinsky closure leaves 0).  FP32-bulk stays within
1e-3 relative L2 of FP64 on the hills."""
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
RAY = dict(damp_opt=2, dampcoef=0.01, zdamp=5000.0)  # s^-1: the explicit relaxation stays stable
DKM = dict(diff_opt=1, kvdif=5.0, damp_opt=1, dampcoef=0.2, zdamp=5000.0)


def _cases():
    return {
        "rayleigh_bubble": (lambda: Case.bubble3d(22, 18, 20, dx=1000.0, ztop=12000.0, dt=6.0, seed=5, u0=6.0,
                                                  bubble_zc=3000.0), dict(RAY, time_step_sound=6), False),
        "rayleigh_symmetric": (lambda: Case.bubble3d(20, 16, 18, dx=1000.0, ztop=12000.0, dt=6.0, seed=7, u0=4.0,
                                                     bubble_zc=3000.0, bc=A.MI_BC_SYMMETRIC),
                               dict(RAY, time_step_sound=6, zdamp=4000.0), False),
        "rayleigh_straka": (lambda: Case.straka(128, 32, dx=200.0, dt=1.0),
                            dict(RAY, zdamp=2000.0, diff_opt=1, km_opt=1, khdif=75.0, kvdif=75.0, pr_inv=1.0), False),
        "rayleigh_hills_spec": (lambda: Case.hills(24, 20, 16, dx=15000.0, dt=90.0, seed=4, bc=A.MI_BC_SPECIFIED),
                                dict(RAY, time_step_sound=4, zdamp=6000.0, dampcoef=0.003, n_moist=2, n_scalars=3,
                                     diff_6th_opt=1),
                                True),
        "dampkm_km1_bubble": (lambda: Case.bubble3d(20, 16, 20, dx=1000.0, dy=700.0, ztop=12000.0, dt=5.0, seed=3,
                                                    bubble_zc=3000.0),
                              dict(DKM, km_opt=1, khdif=50.0, time_step_sound=6, n_moist=1, n_scalars=2), False),
        "dampkm_km4_open": (lambda: Case.bubble3d(24, 20, 16, dx=1000.0, ztop=10000.0, dt=5.0, seed=6,
                                                  bubble_zc=3000.0, bc=A.MI_BC_OPEN),
                            dict(DKM, km_opt=4, time_step_sound=6), False),
        "dampkm_km4_hills_spec": (lambda: Case.hills(24, 20, 16, dx=15000.0, dt=90.0, seed=8, bc=A.MI_BC_SPECIFIED),
                                  dict(DKM, km_opt=4, time_step_sound=4, zdamp=6000.0, n_moist=2, n_scalars=3), True),
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
    # winds that differ from the base profiles, with shear (the Smagorinsky K varies)
    for fid, amp in ((A.MI_U, 2.0), (A.MI_V, -1.5)):
        if g.ny == 1 and fid == A.MI_V:
            continue
        f = gpu.field(fid)
        zz, yy, xx = np.indices(f.shape)
        f = f + amp * np.sin(2 * np.pi * xx / g.nx + 0.3 * zz) * np.cos(2 * np.pi * yy / max(g.ny, 1) + 0.2)
        gpu.set_field(fid, f)
        if orc:
            orc.set_field(fid, f)
    z, y, x = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
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
def test_damping_phases_bitwise(ctx, name):
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[damping] phases {name} grid {case.grid}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    tendf = TENDF if case.grid.ny > 1 else (A.MI_RU_TENDF, A.MI_RW_TENDF, A.MI_T_TENDF)
    ks = KS if prm.damp_opt == 1 else ()
    for step in range(2):
        for d in (gpu, orc):
            d.step_begin()
        for rk in (1, 2, 3):
            for d in (gpu, orc):
                d.rk_prep(rk)
                d.rk_tendency(rk)
            if rk == 1:
                _compare(gpu, orc, ks + tendf, f"step {step} held tendencies")
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
    if prm.damp_opt == 1:
        kmh = gpu.field(A.MI_XKMH)
        assert kmh[-1].max() > kmh[0].max()  # raised near the top


@pytest.mark.parametrize("name", ["rayleigh_bubble", "rayleigh_hills_spec", "dampkm_km4_hills_spec"])
def test_damping_graph_steps_bitwise(ctx, name):
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[damping] graph steps {name}", flush=True)
    for chunk in (1, 4, 5):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE, f"after a chunk of {chunk} steps")


@pytest.mark.parametrize("name", ["rayleigh_hills_spec", "dampkm_km4_hills_spec"])
def test_damping_fp32_close_to_fp64(ctx, name):
    print(f"[damping] {name} FP32-bulk vs FP64, 20 steps", flush=True)
    _, _, g64, _ = _pair(ctx, name, A.MI_FP64)
    _, _, g32, _ = _pair(ctx, name, A.MI_FP32_BULK)
    g64.step(20)
    g32.step(20)
    for n, v in W.error_report(g32, g64, True).items():
        assert v["rel_l2"] <= 1e-3, (n, v)


def test_damping_options_refused(ctx):
    case = Case.bubble3d(16, 12, 10, dx=1000.0, ztop=8000.0, dt=6.0, seed=1)
    for kw, msg in ((dict(damp_opt=4), "damp_opt"), (dict(damp_opt=2, zdamp=0.0), "zdamp")):
        with pytest.raises(Exception, match=msg):
            RK3(ctx, case.grid, case.params(**kw), case)
