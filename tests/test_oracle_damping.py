"""Pins of the oracle's damping layers of the diffusion driver: damp_opt 2, the Rayleigh relaxation of u, v, w and
theta in the top zdamp metres (reference src/dyn/diffusion/rayleigh_damp.hpp:14-259, called last in
DiffusionDriver::dynamics, diffusion_driver.cpp:283-289), and damp_opt 1, the eddy viscosities raised to a damping
profile (cal_dampkm, src/dyn/diffusion/km_kh.hpp:485-568, in DiffusionDriver::prepare).  No fixture of the reference
covers either, so each check is a closed form of the published scheme (WRF's damp_opt 1 and 2):

damp_opt 2, with d = ztop - z the depth of a point below the model top:
 - u = delta over a flat rest state (u_base = 0): ru_tendf = -(c1h mu + c2h) dampcoef sin^2(pi/2 (1 - min(1,
   d / zdamp))) delta on the damped levels and 0 below them; the same for v and for theta = t_base + delta; w = delta
   at every w level of the layer with c1f, c2f;
 - the damped levels of a u, v or theta column are those inside the layer from level 1 up (a layer deeper than
   the model leaves level 0 undamped, w's levels all damped);
 - the relaxation target is the base profile interpolated linearly in height: with u_base = a z^2 + b z on a
   terrain-following grid, u - u00 is the interpolant's closed form between the bracketing base levels.
damp_opt 1 (diff_opt 1), with d_k = (zw(nz) + zw(nz-2) - zw(k) - zw(k-1)) / 2 (zw(-1) = zw(0)), the closed form of
WRF's running depth (half the top layer, then one 1 / rdz = (zw(k+1) - zw(k-1)) / 2 per level below it):
 - km_opt 1: xkmh = max(khdif, cos^2(pi/2 min(1, d / zdamp)) min(dx, dy)^2 / dt dampcoef), xkhh the same with
   khdif / prandtl (the damping value itself is not divided by prandtl);
 - km_opt 4: xkmh = max(Smagorinsky K, the damping value) point by point;
 - specified sides: the columns outside one row in from the relaxation zone keep the closure's K.

Each check was seen to fail on a mistake planted in the oracle: cos for sin in the Rayleigh coefficient, the first
damped level one lower (k1 + 1), the upper bracket for the lower one in the interpolation, the top-layer's half
thickness left out of the depth, max(dx, dy) for min(dx, dy), the damping value divided by prandtl for xkhh, and
the specified-side range widened by one."""
import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case
from oracle_lib import OracleDomain

G = 9.81
DC, ZD = 0.2, 6000.0


def _first_stage(orc):
    orc.step_begin()
    orc.rk_prep(1)
    orc.rk_tendency(1)


def _shape(case, fid):
    g = case.grid
    return A.field_shape(fid, g.nx, g.ny, g.nz)


def _lev(case, name, n):
    return np.array([getattr(case.levels, name)[k] for k in range(n)])


def _coef(d, zd=ZD):
    return np.sin(0.5 * np.pi * (1.0 - np.minimum(d / zd, 1.0))) ** 2


@pytest.mark.parametrize("zd", [ZD, 30000.0])  # a layer inside the model, and one deeper than the model
def test_rayleigh_profile_closed_forms(zd):
    case = Case.rest_state(nx=12, ny=10, nz=30)
    g = case.grid
    orc = OracleDomain(case, case.params(damp_opt=2, dampcoef=DC, zdamp=zd, time_step_sound=4))
    delta = 1.5
    for fid in (A.MI_U, A.MI_V, A.MI_W):
        orc.set_field(fid, np.full(_shape(case, fid), delta))
    tb = case.profile(A.MI_T_INIT)
    orc.set_field(A.MI_T_BASE, tb)
    orc.set_field(A.MI_T, tb[:, None, None] + delta + np.zeros(_shape(case, A.MI_T)))
    _first_stage(orc)
    phb = case.profile(A.MI_PHB)
    ph = orc.field(A.MI_PH)[:, 0, 0]
    zw = (phb + ph) / G
    zm = 0.5 * (zw[:-1] + zw[1:])
    ztop = zw[-1]
    mub = float(case.profile(A.MI_MUB)[0])
    c1h, c2h = _lev(case, "c1h", g.nz), _lev(case, "c2h", g.nz)
    c1f, c2f = _lev(case, "c1f", g.nz + 1), _lev(case, "c2f", g.nz + 1)
    damped = (zm >= ztop - zd) & (np.arange(g.nz) >= 1)  # a layer deeper than the model leaves level 0 alone
    assert 2 < damped.sum() < g.nz - 2 if zd == ZD else damped.sum() == g.nz - 1
    want = np.where(damped, -(c1h * mub + c2h) * DC * _coef(ztop - zm, zd) * delta, 0.0)
    for fid in (A.MI_RU_TENDF, A.MI_RV_TENDF, A.MI_T_TENDF):
        got = orc.field(fid)
        np.testing.assert_allclose(got, np.broadcast_to(want[:, None, None], got.shape), rtol=1e-9, atol=1e-12,
                                   err_msg=A.FIELD_NAMES[fid])
    wlev = zw >= ztop - zd
    want_w = np.where(wlev, -(c1f * mub + c2f) * DC * _coef(ztop - zw, zd) * delta, 0.0)
    got = orc.field(A.MI_RW_TENDF)
    np.testing.assert_allclose(got, np.broadcast_to(want_w[:, None, None], got.shape), rtol=1e-9, atol=1e-12)


def test_rayleigh_interpolates_the_base_profile():
    case = Case.rest_state(nx=12, ny=10, nz=30)
    g = case.grid
    orc = OracleDomain(case, case.params(damp_opt=2, dampcoef=DC, zdamp=ZD, time_step_sound=4))
    phb = case.profile(A.MI_PHB)
    ph = orc.field(A.MI_PH)[:, 0, 0]
    zw = (phb + ph) / G
    zm = 0.5 * (zw[:-1] + zw[1:])
    ztop = zw[-1]
    zb = zm + 100.0  # base heights between the model's levels: the target is interpolated, not read off
    a, b = 2e-7, 1e-3
    orc.set_field(A.MI_Z_BASE, zb)
    orc.set_field(A.MI_U_BASE, a * zb ** 2 + b * zb)
    orc.set_field(A.MI_U, np.zeros(_shape(case, A.MI_U)))
    _first_stage(orc)
    # the piecewise-linear interpolant of a z^2 + b z between the base levels bracketing zm (extrapolated past the
    # top base level)
    k2 = np.clip(np.searchsorted(zb, zm, side="right") - 1, 0, g.nz - 2)
    za, zc = zb[k2], zb[k2 + 1]
    fa, fc = a * za ** 2 + b * za, a * zc ** 2 + b * zc
    u00 = fa + (fc - fa) * (zm - za) / (zc - za)
    assert np.abs(u00 - (a * zm ** 2 + b * zm)).max() > 1e-4  # the interpolation error is visible
    mub = float(case.profile(A.MI_MUB)[0])
    c1h, c2h = _lev(case, "c1h", g.nz), _lev(case, "c2h", g.nz)
    damped = (zm >= ztop - ZD) & (np.arange(g.nz) >= 1)
    want = np.where(damped, (c1h * mub + c2h) * DC * _coef(ztop - zm) * u00, 0.0)
    got = orc.field(A.MI_RU_TENDF)
    np.testing.assert_allclose(got, np.broadcast_to(want[:, None, None], got.shape), rtol=1e-9, atol=1e-12)


def _dampk(case, orc_params, khdif):
    """the closed form of damp_opt 1 on a flat grid: max(khdif, cos^2 kmmax dampcoef) per level"""
    g = case.grid
    phb = case.profile(A.MI_PHB)
    zw = phb / G
    nz = g.nz
    zwm1 = np.concatenate([[zw[0]], zw[:nz - 1]])  # zw(k - 1), zw(0) at k = 0
    d = 0.5 * (zw[nz] + zw[nz - 2] - zw[:nz] - zwm1)
    kmmax = min(g.dx, g.dy) ** 2 / g.dt
    dk = np.cos(0.5 * np.pi * np.minimum(d / ZD, 1.0)) ** 2 * kmmax * DC
    return np.maximum(khdif, dk), dk


def test_dampkm_closed_form_km1():
    case = Case.bubble3d(14, 12, 30, dx=1000.0, dy=700.0, ztop=16000.0, dt=5.0, seed=2, noise_amp=0.0)
    khdif = 40.0
    prm = case.params(diff_opt=1, km_opt=1, khdif=khdif, kvdif=1.0, pr_inv=3.0, damp_opt=1, dampcoef=DC, zdamp=ZD,
                      time_step_sound=4)
    orc = OracleDomain(case, prm)
    orc.set_field(A.MI_PH, np.zeros(_shape(case, A.MI_PH)))
    _first_stage(orc)
    kmh, khh = orc.field(A.MI_XKMH), orc.field(A.MI_XKHH)
    want, dk = _dampk(case, prm, khdif)
    assert (dk > khdif).sum() > 3 and (dk < khdif).sum() > 3
    np.testing.assert_allclose(kmh, np.broadcast_to(want[:, None, None], kmh.shape), rtol=1e-12)
    want_h = np.maximum(khdif * 3.0, dk)
    np.testing.assert_allclose(khh, np.broadcast_to(want_h[:, None, None], khh.shape), rtol=1e-12)


def test_dampkm_with_smagorinsky_is_the_maximum():
    case = Case.bubble3d(16, 14, 24, dx=1000.0, ztop=12000.0, dt=5.0, seed=4, u0=3.0)

    def run(**kw):
        prm = case.params(diff_opt=1, kvdif=1.0, zdamp=ZD, dampcoef=DC, time_step_sound=4, **kw)
        orc = OracleDomain(case, prm)
        zz, yy, xx = np.indices(_shape(case, A.MI_U))
        orc.set_field(A.MI_U, 2.0 * np.sin(2 * np.pi * xx / case.grid.nx) * np.cos(2 * np.pi * yy / case.grid.ny))
        _first_stage(orc)
        return orc.field(A.MI_XKMH), orc.field(A.MI_XKHH)

    smag_m, smag_h = run(km_opt=4, damp_opt=0)
    damp_m, damp_h = run(km_opt=1, khdif=0.0, damp_opt=1)
    both_m, both_h = run(km_opt=4, damp_opt=1)
    assert np.any(damp_m > smag_m) and np.any(damp_m < smag_m)
    assert np.array_equal(both_m, np.maximum(smag_m, damp_m))
    assert np.array_equal(both_h, np.maximum(smag_h, damp_h))


def test_dampkm_specified_range():
    case = Case.hills(20, 18, 16, dx=15000.0, dt=60.0, seed=5, moist=False, bc=A.MI_BC_SPECIFIED)
    khdif = 10.0
    prm = case.params(diff_opt=1, km_opt=1, khdif=khdif, kvdif=1.0, damp_opt=1, dampcoef=DC, zdamp=8000.0,
                      time_step_sound=4)
    orc = OracleDomain(case, prm)
    _first_stage(orc)
    kmh = orc.field(A.MI_XKMH)
    top = kmh[-1]
    inner = top[4:-4, 4:-4]
    assert np.all(inner > khdif)
    outer = np.ones_like(top, bool)
    outer[4:-4, 4:-4] = False
    assert np.all(top[outer] == khdif)
