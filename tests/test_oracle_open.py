"""The oracle's open lateral boundaries (WRF open_xs/xe, open_ys/ye; reference kernels_common.hpp:33-81,
large_step.cpp:152-668, small_step.cpp:151-190, scalars.cpp:42-291) checked against what fixes them
independently of their formulas: the interior of a one-step run does not see the boundary kind; a state at rest
stays at rest; a mirror-symmetric bubble with a tracer stays mirror-symmetric across the open direction (the
two boundaries' terms are mirror images: a one-sided mistake, a wrong neighbour or sign at one edge, breaks it);
the positive-definite tracer stays non-negative."""
import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case
from oracle_lib import OracleDomain

FIELDS = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P)


def test_interior_does_not_see_open_boundaries():
    out = {}
    for bc in (A.MI_BC_PERIODIC, A.MI_BC_OPEN):
        case = Case.bubble3d(84, 84, 10, dx=1000.0, ztop=8000.0, dt=6.0, seed=2, bubble_zc=3000.0, bc=bc)
        o = OracleDomain(case, case.params(time_step_sound=4, damp_opt=3, zdamp=3000.0, diff_6th_opt=1))
        o.step(1)
        out[bc] = {fid: o.field(fid) for fid in FIELDS}
    inner = (slice(None), slice(38, 46), slice(38, 46))
    for fid, a in out[A.MI_BC_PERIODIC].items():
        b = out[A.MI_BC_OPEN][fid]
        assert np.array_equal(a[inner], b[inner]), A.FIELD_NAMES[fid]
        assert not np.array_equal(a, b)  # the edges do differ


@pytest.mark.parametrize("bcx,bcy", [(A.MI_BC_OPEN, A.MI_BC_PERIODIC), (A.MI_BC_SYMMETRIC, A.MI_BC_OPEN)])
def test_rest_state_stays_at_rest(bcx, bcy):
    case = Case.rest_state(nx=16, ny=14, nz=12, bc=bcx, bc_y=bcy)
    o = OracleDomain(case, case.params(time_step_sound=4))
    before = {fid: o.field(fid) for fid in FIELDS}
    o.step(3)
    for fid in (A.MI_U, A.MI_V, A.MI_W):
        assert np.abs(o.field(fid)).max() < 1e-9, A.FIELD_NAMES[fid]
    for fid in (A.MI_T, A.MI_MU):
        assert np.allclose(o.field(fid), before[fid], rtol=0, atol=1e-8), A.FIELD_NAMES[fid]


def _mirror(a, axis, odd):
    m = np.flip(a, axis=axis)
    return -m if odd else m


@pytest.mark.parametrize("direction,cross", [("x", False), ("y", False), ("x", True)])
def test_mirror_symmetry_across_the_open_direction(direction, cross):
    """A bubble, an outflow wind odd in the open direction (out of both boundaries in the lower half, none above
    it; with `cross` also a wind along the boundary, even in the open direction) and a tracer even in it,
    periodic across: after 4 steps every field is its own mirror image (u odd in x, v odd in y) to roundoff and
    the limited tracer stays >= 0.  The calm upper half keeps WRF's one deliberate asymmetry out: rhs_ph gives
    the west lid level no inflow term while the east one has it (large_step.cpp:634-636, 652)."""
    bcx, bcy = (A.MI_BC_OPEN, A.MI_BC_PERIODIC) if direction == "x" else (A.MI_BC_PERIODIC, A.MI_BC_OPEN)
    nx, ny, nz = 28, 24, 12
    case = Case.bubble3d(nx, ny, nz, dx=1000.0, ztop=8000.0, dt=6.0, seed=2, noise_amp=0.0, bubble_zc=3000.0,
                         bubble_radius=3000.0, bc=bcx, bc_y=bcy)
    prm = case.params(time_step_sound=4, damp_opt=3, zdamp=3000.0, n_moist=0, n_scalars=1, scalar_adv_opt=1,
                      diff_6th_opt=1)
    o = OracleDomain(case, prm)
    n = nx if direction == "x" else ny
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    s = (x - nx / 2) if direction == "x" else (y - ny / 2)
    o.set_field(A.MI_Q, (1.0 + 0.5 * np.cos(1.2 * np.pi * s / n)) * np.exp(-((z - 3.0) / 3.0) ** 2))
    prof = np.clip(1.0 - np.arange(nz) / (nz / 2), 0.0, None)[:, None, None]
    if direction == "x":
        xf = np.arange(nx + 1)[None, None, :] - nx / 2
        o.set_field(A.MI_U, 8.0 * np.tanh(xf / 4.0) * prof * np.ones((nz, ny, nx + 1)))
        if cross:
            yv = np.arange(ny + 1)[None, :, None]
            o.set_field(A.MI_V, 4.0 * np.sin(2 * np.pi * yv / ny) * prof * np.ones((nz, ny + 1, nx)))
    else:
        yf = np.arange(ny + 1)[None, :, None] - ny / 2
        o.set_field(A.MI_V, 8.0 * np.tanh(yf / 4.0) * prof * np.ones((nz, ny + 1, nx)))
    o.step(4)
    axis = 2 if direction == "x" else 1
    for fid in FIELDS + (A.MI_Q,):
        a = o.field(fid)
        odd = (fid == A.MI_U and direction == "x") or (fid == A.MI_V and direction == "y")
        scale = max(np.abs(a).max(), 1e-30)
        err = np.abs(a - _mirror(a, axis, odd)).max() / scale
        assert err < 1e-7, f"{A.FIELD_NAMES[fid]}: mirror asymmetry {err:.2e}"
    assert o.field(A.MI_Q).min() >= 0.0
