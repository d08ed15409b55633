"""Pins of the oracle's second-order diffusion along eta with constant K (diff_opt 1, km_opt 1; reference
src/dyn/diffusion/big_step_diffusion.hpp, src/dyn/diffusion_driver.cpp:159-193, 297-327) beyond the reference's
straka golden (test_oracle_golden.py), which is an x-z slice: it pins the x terms of u, w and theta and their
vertical diffusion, not the y terms, v, or the scalars.

On a state at rest with unit map factors and a uniform column mass m (= mub here), one first-stage tendency call
gives closed forms for quadratic fields q = a s^2 along one horizontal direction s, whose centred second difference
is exactly 2 a ds^2:
 - u, w and theta (about t_init) along x or y: 2 a K m (the face mass factors all m);
 - v along x: 2 a K m; v along y: 2 a K, no mass factor (the reference's KHorizDiffV1 follows WRF v4.8.0 lines
   2854-2856 there, big_step_diffusion.hpp:85-87);
 - theta and the scalars take K / prandtl = K pr_inv, the momenta K;
 - the scalars' first stage at rest: q changes by dt/3 * 2 a K pr_inv (their tendency carries muts = m, the update
   divides by it).
The vertical diffusion of v is the mirror of u's (pinned by straka): with the same profiles about their own base
profiles the two tendencies agree.  qv diffuses about qv_base, the other scalars about nothing: a scalar equal to
qv_base's profile changes, qv equal to it does not (a flipped base sign doubles the difference instead).

Each check was seen to fail on a mistake planted in the oracle: the mass factor added to v's y term or dropped from
u's, a wrong column mass in the scalars' horizontal term, xkmh for xkhh in theta's, v_base replaced by u_base, qv's
base left out of its vertical diffusion or subtracted with the wrong sign, and one face coefficient of w off by
0.1 %."""
import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case
from oracle_lib import OracleDomain

K = 150.0


def _domain(khdif, kvdif, pr_inv=3.0, n_moist=0, n_scalars=0):
    case = Case.rest_state(nx=20, ny=18, nz=8)
    prm = case.params(diff_opt=1, km_opt=1, khdif=khdif, kvdif=kvdif, pr_inv=pr_inv, n_moist=n_moist,
                      n_scalars=n_scalars, time_step_sound=4)
    orc = OracleDomain(case, prm)
    return case, orc


def _tendencies(orc):
    orc.step_begin()
    orc.rk_prep(1)
    orc.rk_tendency(1)
    return {f: orc.field(f) for f in (A.MI_RU_TENDF, A.MI_RV_TENDF, A.MI_RW_TENDF, A.MI_T_TENDF)}


def _quadratic(case, fid, axis, a):
    """a s^2 along x (axis 2) or y (axis 1) at the field's points, s in metres"""
    shape = A.field_shape(fid, case.grid.nx, case.grid.ny, case.grid.nz)
    idx = np.indices(shape)[axis].astype(float)
    ds = case.grid.dx if axis == 2 else case.grid.dy
    return a * (idx * ds) ** 2


INNER = (slice(None), slice(3, -3), slice(3, -3))  # away from the periodic wrap of the quadratic


def _mub(case):
    return float(case.profile(A.MI_MUB)[0])


@pytest.mark.parametrize("axis", [2, 1])
def test_momentum_and_theta_closed_forms(axis):
    a = 1e-9
    case, orc = _domain(K, 0.0)
    m = _mub(case)
    t0 = case.field(A.MI_T)
    orc.set_field(A.MI_U, _quadratic(case, A.MI_U, axis, a))
    orc.set_field(A.MI_V, _quadratic(case, A.MI_V, axis, a))
    orc.set_field(A.MI_W, _quadratic(case, A.MI_W, axis, a))
    orc.set_field(A.MI_T, t0 + _quadratic(case, A.MI_T, axis, a))
    t = _tendencies(orc)
    ru = t[A.MI_RU_TENDF][INNER]
    rv = t[A.MI_RV_TENDF][INNER]
    rw = t[A.MI_RW_TENDF][1:-1, 3:-3, 3:-3]
    tt = t[A.MI_T_TENDF][INNER]
    np.testing.assert_allclose(ru, 2 * a * K * m, rtol=1e-9)
    np.testing.assert_allclose(rv, 2 * a * K * (m if axis == 2 else 1.0), rtol=1e-9)
    np.testing.assert_allclose(rw, 2 * a * K * m, rtol=1e-9)
    np.testing.assert_allclose(tt, 2 * a * K * 3.0 * m, rtol=1e-7)  # theta: K / prandtl (the base state cancels)


def test_vertical_v_mirrors_u():
    case, orc = _domain(0.0, 20.0)
    nz = case.grid.nz
    prof = 3.0 * np.sin(0.7 * np.arange(nz)) + 0.2 * np.arange(nz) ** 2
    ub, vb = 0.5 * np.arange(nz), -0.3 * np.arange(nz) ** 1.5
    u = np.broadcast_to((prof + ub)[:, None, None], A.field_shape(A.MI_U, case.grid.nx, case.grid.ny, nz))
    v = np.broadcast_to((prof + vb)[:, None, None], A.field_shape(A.MI_V, case.grid.nx, case.grid.ny, nz))
    orc.set_field(A.MI_U, u)
    orc.set_field(A.MI_V, v)
    orc.set_field(A.MI_U_BASE, ub)
    orc.set_field(A.MI_V_BASE, vb)
    t = _tendencies(orc)
    ru, rv = t[A.MI_RU_TENDF][:, 3:15, 3:17], t[A.MI_RV_TENDF][:, 3:15, 3:17]
    assert np.abs(ru).max() > 0
    np.testing.assert_allclose(rv, ru, rtol=1e-9, atol=1e-12 * np.abs(ru).max())
    assert np.all(ru[-1] == 0)  # vertical_diffusion_u leaves the top level alone


def _scalar_first_stage(orc):
    orc.step_begin()
    orc.rk_prep(1)
    orc.rk_tendency(1)
    orc.small_steps(1)
    orc.scalar_transport(1)


def test_scalar_closed_form_and_qv_base():
    case, orc = _domain(K, 0.0, n_moist=1, n_scalars=2)
    a = 1e-12
    g = case.grid
    q = 1e-3 + _quadratic(case, A.MI_Q, 2, a)
    orc.set_field(A.MI_Q + 1, q)
    _scalar_first_stage(orc)
    dq = orc.field(A.MI_Q + 1)[INNER] - q[INNER]
    np.testing.assert_allclose(dq, g.dt / 3.0 * 2 * a * K * 3.0, rtol=1e-6)

    # vertical: qv about qv_base, a tracer about nothing
    case, orc = _domain(0.0, 50.0, n_moist=1, n_scalars=2)
    nz = case.grid.nz
    base = 0.012 * np.exp(-np.arange(nz) / 3.0)
    shape = A.field_shape(A.MI_Q, case.grid.nx, case.grid.ny, nz)
    prof = np.broadcast_to(base[:, None, None], shape).copy()
    orc.set_field(A.MI_QV_BASE, base)
    orc.set_field(A.MI_Q, prof)
    orc.set_field(A.MI_Q + 1, prof)
    _scalar_first_stage(orc)
    dqv = orc.field(A.MI_Q) - prof
    dtr = orc.field(A.MI_Q + 1) - prof
    assert np.abs(dtr).max() > 1e-9
    assert np.abs(dqv).max() < 1e-6 * np.abs(dtr).max()
