"""Checkpoint and resume of a domain's carried state (no counterpart in the reference's src/dyn; its restart test,
tests/robustness/run_test.sh:40-51, resumes from NetCDF restart files): a resumed forecast equals the
uninterrupted one bit for bit, in a fresh domain and after rewinding the same one, in FP64 and FP32-bulk, with
specified boundaries and scalars; in FP64 the resumed forecast is the oracle's uninterrupted one bit for bit."""
import numpy as np
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


def _fields(d, fids):
    return {fid: d.field(fid) for fid in fids}


def _same(a, b, where):
    for fid in a:
        assert bitwise_equal(a[fid], b[fid]), f"{where}: {A.FIELD_NAMES[fid]}: {first_diff(a[fid], b[fid])}"


@pytest.mark.parametrize("precision", [A.MI_FP64, A.MI_FP32_BULK])
def test_resume_is_bitwise(ctx, precision):
    case = Case.hills(24, 20, 18, dx=15000.0, dt=90.0, seed=3, bc=A.MI_BC_SPECIFIED)
    prm = case.params(precision=precision, time_step_sound=4, damp_opt=3, n_moist=3, n_scalars=4, moist_adv_opt=1)
    bdy = drifting_boundary(case, seed=11, q_species=(0,))

    def make():
        d = RK3(ctx, case.grid, prm, case)
        d.load_boundary(bdy)
        return d

    print(f"[checkpoint] precision {precision}", flush=True)
    fids = STATE + tuple(A.MI_Q + n for n in range(prm.n_scalars))
    d1 = make()
    d1.step(3)
    buf = d1.save_state()
    d1.step(4)
    want = _fields(d1, fids)
    t_want = d1.bdy_time
    d2 = make()  # a fresh domain with the same inputs, resumed
    d2.load_state(buf)
    assert d2.bdy_time == 3 * case.grid.dt
    d2.step(4)
    _same(_fields(d2, fids), want, "resumed in a fresh domain")
    assert d2.bdy_time == t_want
    d1.load_state(buf)  # rewind the same domain (its scalar buffers have swapped since the save)
    d1.step(4)
    _same(_fields(d1, fids), want, "rewound")
    if precision == A.MI_FP64:  # the resumed forecast is the oracle's 7 uninterrupted steps
        orc = OracleDomain(case, prm)
        orc.load_boundary(bdy)
        orc.step(7)
        _same(_fields(orc, fids), want, "against the oracle")


def test_rejects_foreign_state(ctx):
    a = Case.bubble3d(20, 16, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=5)
    b = Case.bubble3d(20, 16, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=5)
    da = RK3(ctx, a.grid, a.params(time_step_sound=6), a)
    db = RK3(ctx, b.grid, b.params(time_step_sound=6), b)
    d32 = RK3(ctx, a.grid, a.params(time_step_sound=6, precision=A.MI_FP32_BULK), a)
    buf = da.save_state()
    with pytest.raises(A.MiError):
        db.load_state(buf)  # another nz
    with pytest.raises(A.MiError):
        d32.load_state(buf)  # another storage precision
    with pytest.raises(A.MiError):
        da.load_state(buf[:100])  # truncated
    bad = buf.copy()
    bad[0] ^= 0xFF
    with pytest.raises(A.MiError):
        da.load_state(bad)
    da.load_state(buf)  # its own state is accepted
