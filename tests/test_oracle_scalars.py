"""The oracle's moisture coupling and scalar transport, pinned by the invariants the reference's own tests
check (tests/unit/test_scalars.cpp:66-93, tests/unit/test_moist.cpp:79-85):

- a constant tracer stays constant (to 1e-12) under advect_scalar and advect_scalar_pd;
- the positive-definite limiter keeps a blob non-negative (>= -1e-12 of its maximum), and without it the
  5th/3rd-order fluxes undershoot (< -1e-4 of the maximum), so the option is really applied;
- the tracer mass sum (c1h (mub + mu) + c2h) q (-dnw) is conserved to 1e-13 on a closed domain;
- a moist species identically zero leaves the dynamics equal to the dry run (calc_cq then gives
  cqu = cqv = cq1 = 1, cq2 = 0 exactly).
"""
import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case
from oracle_lib import OracleDomain, olib


@pytest.fixture(scope="module", autouse=True)
def threads():
    olib().orc_set_threads(8)


def _levels(case):
    lv, nz = case.levels, case.grid.nz
    return (np.array([lv.c1h[k] for k in range(nz)]), np.array([lv.c2h[k] for k in range(nz)]),
            np.array([lv.dnw[k] for k in range(nz)]))


def tracer_mass(o, case, q):
    c1h, c2h, dnw = _levels(case)
    mu = o.field(A.MI_MU)[0]
    mub = case.profile(A.MI_MUB)[0]
    m = c1h[:, None, None] * (mub + mu)[None] + c2h[:, None, None]
    return float(np.sum(m * q * (-dnw)[:, None, None]))


def blob(case, xc, zc, radius):
    g = case.grid
    z = (np.arange(g.nz) + 0.5) * (case.spec.ztop / g.nz)
    x = (np.arange(g.nx) + 0.5) * g.dx
    r = np.sqrt(((x[None, :] - xc) / radius) ** 2 + ((z[:, None] - zc) / radius) ** 2)
    return np.where(r < 1, np.cos(0.5 * np.pi * r) ** 2, 0.0)[:, None, :] * np.ones((1, g.ny, 1))


@pytest.mark.parametrize("pd", [0, 1])
def test_constant_tracer_positivity_and_mass(pd):
    case = Case.warm_bubble()  # x-z, symmetric walls, a rising bubble
    g = case.grid
    o = OracleDomain(case, case.params(n_scalars=2, scalar_adv_opt=pd))
    b0 = blob(case, 10000.0, 3000.0, 1500.0)
    o.set_field(A.MI_Q + 0, np.ones(A.field_shape(A.MI_Q, g.nx, g.ny, g.nz)))
    o.set_field(A.MI_Q + 1, b0)
    m0 = tracer_mass(o, case, b0)
    o.step(200)
    one, b = o.field(A.MI_Q + 0), o.field(A.MI_Q + 1)
    assert np.abs(o.field(A.MI_W)).max() > 1.0  # the flow is not at rest
    assert np.abs(one - 1.0).max() < 1e-12
    if pd:
        assert b.min() >= -1e-12 * b0.max()
    else:
        assert b.min() < -1e-4 * b0.max()  # the limiter is what removes the undershoot
    assert abs(tracer_mass(o, case, b) - m0) / m0 < 1e-13


def test_periodic_3d_mass_and_positivity():
    case = Case.bubble3d(16, 14, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=4, bubble_zc=3000.0)
    g = case.grid
    o = OracleDomain(case, case.params(time_step_sound=6, n_moist=1, n_scalars=2, moist_adv_opt=1, scalar_adv_opt=1))
    rng = np.random.default_rng(3)
    shp = A.field_shape(A.MI_Q, g.nx, g.ny, g.nz)
    qv = 0.01 * np.exp(-np.arange(g.nz) / 4.0)[:, None, None] * (1 + 0.1 * rng.random(shp))
    tr = (rng.random(shp) > 0.7).astype(float)  # sharp-edged tracer
    o.set_field(A.MI_Q + 0, qv)
    o.set_field(A.MI_Q + 1, tr)
    m0 = [tracer_mass(o, case, qv), tracer_mass(o, case, tr)]
    o.step(10)
    for n, q0 in ((0, qv), (1, tr)):
        q = o.field(A.MI_Q + n)
        assert q.min() >= -1e-12 * q0.max()
        assert abs(tracer_mass(o, case, q) - m0[n]) / m0[n] < 1e-13


def test_zero_moisture_equals_dry():
    case = Case.bubble3d(16, 14, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=4, bubble_zc=3000.0)
    g = case.grid
    dry = OracleDomain(case, case.params(time_step_sound=6))
    wet = OracleDomain(case, case.params(time_step_sound=6, n_moist=1, n_scalars=1))
    wet.set_field(A.MI_Q, np.zeros(A.field_shape(A.MI_Q, g.nx, g.ny, g.nz)))
    dry.step(5)
    wet.step(5)
    for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL):
        assert np.array_equal(dry.field(fid), wet.field(fid)), A.FIELD_NAMES[fid]
    assert np.array_equal(wet.field(A.MI_CQU), np.ones(A.field_shape(A.MI_CQU, g.nx, g.ny, g.nz)))


def test_moisture_changes_the_dynamics():
    """cqu, cqv, cqw enter the pressure-gradient terms: a moist run departs from the dry one."""
    case = Case.bubble3d(16, 14, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=4, bubble_zc=3000.0)
    g = case.grid
    dry = OracleDomain(case, case.params(time_step_sound=6))
    wet = OracleDomain(case, case.params(time_step_sound=6, n_moist=1, n_scalars=1))
    qv = 0.01 * np.exp(-np.arange(g.nz) / 4.0)[:, None, None] * np.ones(A.field_shape(A.MI_Q, g.nx, g.ny, g.nz))
    wet.set_field(A.MI_Q, qv)
    dry.step(2)
    wet.step(2)
    cqw = wet.field(A.MI_CQW)
    assert np.all(cqw[1:-1] > 0) and np.all(cqw[0] == 0)
    assert not np.array_equal(dry.field(A.MI_W), wet.field(A.MI_W))


def test_uniform_tracer_in_uniform_flow():
    """A uniform tracer carried by a uniform 3-D flow stays uniform through the PD stage as well."""
    case = Case.rest_state(nx=12, ny=10, nz=10, u0=5.0)
    g = case.grid
    o = OracleDomain(case, case.params(time_step_sound=4, n_scalars=1, scalar_adv_opt=1))
    o.set_field(A.MI_Q, 0.5 * np.ones(A.field_shape(A.MI_Q, g.nx, g.ny, g.nz)))
    o.step(3)
    assert np.abs(o.field(A.MI_Q) - 0.5).max() < 1e-13
