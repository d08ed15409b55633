"""FP32-bulk mode (MI_FP32_BULK): the 3-D bulk fields stored in FP32, with FP64 kept for the base state, the
column masses, php, the advance_w tridiagonal solve and calc_p_rho.  Checked against the FP64 path on the
same inputs: relative L2 error ||a - b|| / ||b|| and max |a - b| per variable, as the reference's FP32
study (docs/FP32_ACCURACY.md:24-35); the bar is 1e-3 in relative L2 for every variable."""
import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case, Context, RK3

pytestmark = pytest.mark.gpu

VARS = {"U": A.MI_U, "V": A.MI_V, "W": A.MI_W, "T": A.MI_T, "PH": A.MI_PH, "MU": A.MI_MU}


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def errors(d32, d64, extra=()):
    out = {}
    for name, fid in list(VARS.items()) + list(extra):
        a, b = d32.field(fid), d64.field(fid)
        out[name] = (rel_l2(a, b), float(np.abs(a - b).max()))
    return out


def test_fp32_bubble64_100_steps(ctx):
    case = Case.bubble3d(64, 64, 40, dx=1000.0, ztop=20000.0, dt=6.0, seed=1, noise_amp=0.05)
    kw = dict(time_step_sound=6, damp_opt=3, zdamp=5000.0, dampcoef=0.2)
    print("[fp32] bubble 64x64x40 100 steps", flush=True)
    d64 = RK3(ctx, case.grid, case.params(**kw), case)
    d32 = RK3(ctx, case.grid, case.params(precision=A.MI_FP32_BULK, **kw), case)
    d64.step(100)
    d32.step(100)
    assert d32.check_finite() == -1
    err = errors(d32, d64)
    print(err, flush=True)
    for name, (l2, mx) in err.items():
        assert l2 < 1e-3, (name, l2, mx)


def test_fp32_hills_moist_one_hour(ctx):
    """The real-data-shaped domain at 15 km (dt 90 s, 4 acoustic steps): 40 steps = 1 simulated hour."""
    case = Case.hills(60, 48, 30, dx=15000.0, dt=90.0, seed=2)
    kw = dict(time_step_sound=4, damp_opt=3, zdamp=5000.0, diff_6th_opt=2)
    print("[fp32] hills 60x48x30 1 h", flush=True)
    d64 = RK3(ctx, case.grid, case.params(**kw), case)
    d32 = RK3(ctx, case.grid, case.params(precision=A.MI_FP32_BULK, **kw), case)
    d64.step(40)
    d32.step(40)
    assert d32.check_finite() == -1
    err = errors(d32, d64, extra=[("QVAPOR", A.MI_Q)])
    print(err, flush=True)
    for name, (l2, mx) in err.items():
        assert l2 < 1e-3, (name, l2, mx)


def test_fp32_phases_and_graph_agree(ctx):
    """The FP32 step through the phase entry points equals the graph-replayed FP32 step bit for bit."""
    case = Case.bubble3d(20, 16, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=7, bubble_zc=3000.0)
    prm = case.params(precision=A.MI_FP32_BULK, time_step_sound=6, n_moist=1, n_scalars=2)
    a = RK3(ctx, case.grid, prm, case)
    b = RK3(ctx, case.grid, prm, case)
    g = case.grid
    q = 0.01 * np.ones(A.field_shape(A.MI_Q, g.nx, g.ny, g.nz))
    for d in (a, b):
        d.set_field(A.MI_Q, q)
    a.step(1)
    b.step_begin()
    for rk in (1, 2, 3):
        b.rk_prep(rk)
        b.rk_tendency(rk)
        b.small_steps(rk)
        b.scalar_transport(rk)
        b.calc_p_rho_phi(rk)
    for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL, A.MI_Q):
        assert np.array_equal(a.field(fid), b.field(fid)), A.FIELD_NAMES[fid]
