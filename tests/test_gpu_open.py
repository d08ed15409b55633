"""Open lateral boundaries (WRF open_xs/xe, open_ys/ye: the radiation faces, the inflow terms of every advected
field, the degraded stencils and diffusion limits; reference kernels_common.hpp:33-81, large_step.cpp:152-668,
small_step.cpp:151-190, 279-290, scalars.cpp:42-291, sixth_order.hpp:179-204) on the GPU: bit for bit against
the oracle phase by phase and over graph-replayed steps, with winds through the boundaries in both directions,
moisture, the positive-definite limiter and diffusion; FP32-bulk close to FP64.  The oracle's open path is
checked against invariants that fix it (mirror symmetry, the rest state, the interior's independence of the
boundary kind) in test_oracle_open.py."""
import numpy as np
import pytest

from conftest import bitwise_equal, first_diff
from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Batch, Case, Context, RK3
from oracle_lib import OracleDomain

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)
PREP = (A.MI_WW, A.MI_ALT, A.MI_PHP)
TEND = (A.MI_RU_TEND, A.MI_RV_TEND, A.MI_RW_TEND, A.MI_T_TEND, A.MI_PH_TEND)
STAGE = STATE + (A.MI_RU_M, A.MI_RV_M, A.MI_WW_M, A.MI_MUTS, A.MI_WW)

O, P, S = A.MI_BC_OPEN, A.MI_BC_PERIODIC, A.MI_BC_SYMMETRIC


def _cases():
    """name -> (case, params, wind): `wind` names the components set to a flow through the open boundaries,
    inward at one edge and outward at the other and varying along them, so both signs of the radiation and
    inflow terms occur (none across a symmetric wall, whose faces carry no normal wind)"""
    bub = dict(dx=1000.0, ztop=8000.0, dt=6.0, bubble_zc=3000.0)
    return {
        "x_bubble": (lambda: Case.bubble3d(20, 16, 14, seed=5, bc=O, bc_y=P, **bub),
                     dict(time_step_sound=6, damp_opt=3, zdamp=3000.0, n_moist=0, n_scalars=2, scalar_adv_opt=1,
                          diff_6th_opt=1), "uv"),
        "y_bubble": (lambda: Case.bubble3d(16, 20, 14, seed=6, bc=S, bc_y=O, **bub),
                     dict(time_step_sound=6, damp_opt=3, zdamp=3000.0, n_moist=0, n_scalars=1, diff_6th_opt=2), "v"),
        "xy_hills_moist": (lambda: Case.hills(24, 20, 18, dx=15000.0, dt=90.0, seed=3, bc=O, bc_y=O),
                           dict(time_step_sound=4, damp_opt=3, n_moist=3, n_scalars=4, diff_6th_opt=2), ""),
        "xy_hills_plain": (lambda: Case.hills(22, 21, 16, dx=15000.0, dt=90.0, seed=4, bc=O, bc_y=O),
                           dict(time_step_sound=4, damp_opt=3, n_moist=2, n_scalars=3, moist_adv_opt=0,
                                scalar_adv_opt=1, diff_6th_opt=1, h_adv_order=6, v_adv_order=4), ""),
        "slice": (lambda: Case.bubble3d(40, 1, 20, seed=7, bc=O, bc_y=P, **bub),
                  dict(time_step_sound=6, damp_opt=3, zdamp=3000.0, n_moist=0, n_scalars=1, scalar_adv_opt=1), "u"),
    }


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _pair(ctx, name, precision=A.MI_FP64):
    make, kw, wind = _cases()[name]
    case = make()
    prm = case.params(precision=precision, **kw)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    g = case.grid
    z, y, x = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    for n in range(prm.n_moist, prm.n_scalars):  # a passive tracer: a smooth blob, nonzero at the edges
        tr = 1.0 + np.cos(0.4 * x + 0.3 * y + n) * np.exp(-0.1 * z)
        gpu.set_field(A.MI_Q + n, tr)
        orc.set_field(A.MI_Q + n, tr)
    if wind:
        prof = np.exp(-((np.arange(g.nz) - g.nz / 3) / (g.nz / 3)) ** 2)[:, None, None]
        yu = np.arange(g.ny)[None, :, None]
        u = (6.0 + 4.0 * np.cos(2 * np.pi * yu / max(g.ny, 2))) * prof * np.ones((g.nz, g.ny, g.nx + 1))
        xv = np.arange(g.nx)[None, None, :]
        v = (-5.0 + 3.0 * np.sin(2 * np.pi * xv / g.nx)) * prof * np.ones((g.nz, g.ny + 1, g.nx))
        for d in (gpu, orc):
            if "u" in wind:
                d.set_field(A.MI_U, u)
            if "v" in wind:
                d.set_field(A.MI_V, v)
    return case, prm, gpu, orc


def _compare(gpu, orc, fids, where):
    for fid in fids:
        a, b = gpu.field(fid), orc.field(fid)
        assert bitwise_equal(a, b), f"{where}: {A.FIELD_NAMES[fid]}: {first_diff(a, b)}"


@pytest.mark.parametrize("name", list(_cases()))
def test_open_phases_bitwise(ctx, name):
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[open] phases {name} grid {case.grid}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    for step in range(2):
        gpu.step_begin()
        orc.step_begin()
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
        _compare(gpu, orc, STATE + qs, f"step {step} end")


@pytest.mark.parametrize("name", ["x_bubble", "xy_hills_moist", "slice"])
def test_open_graph_steps_bitwise(ctx, name):
    """mi_step (graph-replayed) over 12 steps against the oracle"""
    case, prm, gpu, orc = _pair(ctx, name)
    print(f"[open] graph steps {name}", flush=True)
    qs = tuple(A.MI_Q + n for n in range(prm.n_scalars))
    for chunk in (1, 5, 6):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE + qs, f"after a chunk of {chunk} steps")
    assert gpu.check_finite() == -1


def test_open_fp32_close(ctx):
    """FP32-bulk against FP64 with both directions open, 40 steps (1 simulated hour at 15 km): relative L2
    below 1e-3 for every prognostic variable and the vapour"""
    case = Case.hills(48, 40, 24, dx=15000.0, dt=90.0, seed=2, bc=O, bc_y=O)
    kw = dict(time_step_sound=4, damp_opt=3, zdamp=5000.0, diff_6th_opt=2)
    print("[open] fp32 hills 48x40x24 1 h", flush=True)
    d64 = RK3(ctx, case.grid, case.params(**kw), case)
    d32 = RK3(ctx, case.grid, case.params(precision=A.MI_FP32_BULK, **kw), case)
    d64.step(40)
    d32.step(40)
    assert d32.check_finite() == -1
    for name, fid in (("U", A.MI_U), ("V", A.MI_V), ("W", A.MI_W), ("T", A.MI_T), ("PH", A.MI_PH), ("MU", A.MI_MU),
                      ("QVAPOR", A.MI_Q)):
        a, b = d32.field(fid), d64.field(fid)
        l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        print(name, l2, flush=True)
        assert l2 < 1e-3, (name, l2)


def test_open_domains_in_a_batch(ctx):
    """Open-boundary domains of different sizes and directions stepped together in one graph per batch step
    (mi_batch_step) end bit-identical to the oracle"""
    print("[open] batch", flush=True)
    pairs = [_pair(ctx, name) for name in ("x_bubble", "xy_hills_moist", "slice")]
    b = Batch(ctx, [p[2] for p in pairs])
    b.step(3)
    b.sync()
    for case, prm, gpu, orc in pairs:
        orc.step(3)
        _compare(gpu, orc, STATE + tuple(A.MI_Q + n for n in range(prm.n_scalars)), f"batch member {case.grid}")
