"""FP64 parity of the HIP dynamics step with the CPU oracle: bit for bit, phase by phase and over many
steps.  The oracle itself is pinned to the reference's golden files (test_oracle_golden.py)."""
import sys

import numpy as np
import pytest

from conftest import bitwise_equal, first_diff
from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case, Context, RK3
from oracle_lib import OracleDomain

pytestmark = pytest.mark.gpu

STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)
PREP = (A.MI_WW, A.MI_ALT, A.MI_PHP)
TEND = (A.MI_RU_TEND, A.MI_RV_TEND, A.MI_RW_TEND, A.MI_T_TEND, A.MI_PH_TEND)
STAGE = STATE + (A.MI_RU_M, A.MI_RV_M, A.MI_WW_M, A.MI_MUTS, A.MI_WW)
CQ = (A.MI_CQU, A.MI_CQV, A.MI_CQW)


def _cases():
    return {
        "igw": (lambda: Case.igw(), {}),
        "warm_bubble": (lambda: Case.warm_bubble(), {}),
        "rest_u5": (lambda: Case.rest_state(u0=5.0), {}),
        "bubble3d_damp": (lambda: Case.bubble3d(24, 20, 16, dx=1000.0, ztop=8000.0, dt=6.0, seed=7, bubble_zc=3000.0),
                          dict(time_step_sound=6, damp_opt=3, zdamp=3000.0)),
        "bubble3d_h6v4_lid": (lambda: Case.bubble3d(20, 18, 12, dx=1000.0, ztop=6000.0, dt=6.0, seed=3, bubble_zc=2000.0,
                                                    bc=A.MI_BC_SYMMETRIC),
                              dict(time_step_sound=4, h_adv_order=6, v_adv_order=4, top_lid=1, w_damping=1)),
    }


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _compare(gpu, orc, fids, where):
    for fid in fids:
        a, b = gpu.field(fid), orc.field(fid)
        assert bitwise_equal(a, b), f"{where}: {A.FIELD_NAMES[fid]}: {first_diff(a, b)}"


@pytest.mark.parametrize("name", list(_cases()))
def test_phases_bitwise(ctx, name):
    make, kw = _cases()[name]
    case = make()
    prm = case.params(**kw)
    print(f"[parity] phases {name} grid {case.grid}", flush=True)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    assert gpu.n_sound == orc.n_sound
    _compare(gpu, orc, STATE, "initial state")
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
            gpu.calc_p_rho_phi(rk)
            orc.calc_p_rho_phi(rk)
            _compare(gpu, orc, (A.MI_P, A.MI_AL), f"step {step} rk {rk} calc_p_rho_phi")


@pytest.mark.parametrize("name", ["igw", "bubble3d_damp", "bubble3d_h6v4_lid"])
def test_graph_steps_bitwise(ctx, name):
    """mi_step (the step captured as a HIP graph and replayed) against the oracle over 12 steps"""
    make, kw = _cases()[name]
    case = make()
    prm = case.params(**kw)
    print(f"[parity] graph steps {name}", flush=True)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    for chunk in (1, 5, 6):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE, f"after a chunk of {chunk} steps")


def test_bubble64_100_steps_bitwise(ctx):
    """The 64x64x40 warm-bubble workload (dt 6 s, 6 acoustic steps, damp_opt 3), 100 steps: the HIP
    path and the oracle agree in every bit of every prognostic field."""
    case = Case.bubble3d(64, 64, 40, dx=1000.0, ztop=20000.0, dt=6.0, seed=1, noise_amp=0.05)
    prm = case.params(time_step_sound=6, damp_opt=3, zdamp=5000.0, dampcoef=0.2)
    print("[parity] bubble 64x64x40 100 steps", flush=True)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    for chunk in (10, 40, 50):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE, f"bubble64 after +{chunk} steps")
        sys.stdout.flush()
    assert gpu.check_finite() == -1
    w = gpu.field(A.MI_W)
    assert np.abs(w).max() > 1.0  # the bubble rises: the comparison is not of a state at rest


@pytest.mark.parametrize("shape", range(13))
def test_every_block_shape_bitwise(ctx, shape):
    """The column kernels give the same bits for every block shape the autotuner may choose."""
    case = Case.bubble3d(24, 20, 16, dx=1000.0, ztop=8000.0, dt=6.0, seed=7, bubble_zc=3000.0)
    prm = case.params(time_step_sound=6, damp_opt=3, zdamp=3000.0)
    print(f"[parity] block shape {shape}", flush=True)
    gpu = RK3(ctx, case.grid, prm, case)
    # the substep kernel needs a thread per level (16 + 1 here): shapes with fewer keep its default
    threads_per_column = [16, 32, 16, 32, 64, 128, 256, 4, 8, 16, 8, 64, 64]
    acoustic = shape if threads_per_column[shape] >= 17 else gpu.tuning[2]
    gpu.tuning = [shape, shape, acoustic, shape]
    assert gpu.tuning == [shape, shape, acoustic, shape]
    orc = OracleDomain(case, prm)
    gpu.step(3)
    orc.step(3)
    _compare(gpu, orc, STATE, f"shape {shape}")


def test_autotune_keeps_the_trajectory(ctx):
    """Autotuning steps the domain to time the shapes, then restores its state: a tuned run equals the oracle."""
    case = Case.bubble3d(32, 24, 20, dx=1000.0, ztop=10000.0, dt=6.0, seed=5, bubble_zc=3000.0)
    prm = case.params(time_step_sound=6, damp_opt=3, zdamp=4000.0)
    print("[parity] autotune", flush=True)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    gpu.step(2)
    orc.step(2)
    shapes = gpu.autotune(1)
    assert all(0 <= s < 14 for s in shapes)  # 13: the substep kernel's k-march
    gpu.step(4)
    orc.step(4)
    _compare(gpu, orc, STATE, f"after autotune {shapes}")


def _moist(name):
    """A moist case: two moist species (PD at the last stage) and a passive tracer (plain advection), a
    held tendency on each and a microphysics heating field, so every moist term of the step is non-zero."""
    if name == "moist3d":
        case = Case.bubble3d(20, 16, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=9, bubble_zc=3000.0)
        kw = dict(time_step_sound=6, damp_opt=3, zdamp=3000.0)
    else:  # x-z slice with symmetric walls
        case = Case.bubble3d(40, 1, 20, dx=500.0, ztop=8000.0, dt=3.0, seed=2, bubble_zc=2500.0, bc=A.MI_BC_SYMMETRIC)
        kw = dict(time_step_sound=6)
    prm = case.params(n_moist=2, n_scalars=3, moist_adv_opt=1, scalar_adv_opt=0, **kw)
    g = case.grid
    rng = np.random.default_rng(11)
    shp = A.field_shape(A.MI_Q, g.nx, g.ny, g.nz)
    prof = np.exp(-np.arange(g.nz) / 5.0)[:, None, None]
    fields = {
        A.MI_Q + 0: 0.012 * prof * (1 + 0.2 * rng.random(shp)),
        A.MI_Q + 1: 1e-3 * (rng.random(shp) > 0.8),
        A.MI_Q + 2: rng.random(shp),
        A.MI_Q_TENDF + 0: 1e-4 * rng.standard_normal(shp),
        A.MI_Q_TENDF + 1: 1e-5 * rng.standard_normal(shp),
        A.MI_Q_TENDF + 2: 1e-3 * rng.standard_normal(shp),
        A.MI_H_DIABATIC: 1e-3 * rng.standard_normal(shp),
    }
    return case, prm, fields


def _moist_pair(ctx, name):
    case, prm, fields = _moist(name)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    for fid, a in fields.items():
        gpu.set_field(fid, a)
        orc.set_field(fid, a)
    return gpu, orc


@pytest.mark.parametrize("name", ["moist3d", "moist_xz"])
def test_moist_phases_bitwise(ctx, name):
    """calc_cq, the moist terms of the tendencies and acoustic steps, and the scalar transport (plain and
    positive-definite), phase by phase over two steps."""
    print(f"[parity] moist phases {name}", flush=True)
    gpu, orc = _moist_pair(ctx, name)
    QS = tuple(A.MI_Q + n for n in range(3))
    for step in range(2):
        gpu.step_begin()
        orc.step_begin()
        for rk in (1, 2, 3):
            gpu.rk_prep(rk)
            orc.rk_prep(rk)
            _compare(gpu, orc, PREP + CQ, f"step {step} rk {rk} prep")
            gpu.rk_tendency(rk)
            orc.rk_tendency(rk)
            _compare(gpu, orc, TEND, f"step {step} rk {rk} tendencies")
            gpu.small_steps(rk)
            orc.small_steps(rk)
            _compare(gpu, orc, STAGE, f"step {step} rk {rk} acoustic steps")
            gpu.scalar_transport(rk)
            orc.scalar_transport(rk)
            _compare(gpu, orc, QS, f"step {step} rk {rk} scalar transport")
            gpu.calc_p_rho_phi(rk)
            orc.calc_p_rho_phi(rk)
            _compare(gpu, orc, (A.MI_P, A.MI_AL), f"step {step} rk {rk} calc_p_rho_phi")


def test_moist_graph_steps_bitwise(ctx):
    """The replayed step graphs (one per scalar-buffer parity) over 11 steps, moist, with the oracle."""
    print("[parity] moist graph steps", flush=True)
    gpu, orc = _moist_pair(ctx, "moist3d")
    for chunk in (1, 3, 7):
        gpu.step(chunk)
        orc.step(chunk)
        _compare(gpu, orc, STATE + tuple(A.MI_Q + n for n in range(3)), f"moist after a chunk of {chunk} steps")


def _mapped_case():
    """Map factors varying in [0.95, 1.05] (a Lambert-like pattern at every staggering), Coriolis and grid
    rotation fields, sixth-order diffusion (monotonic) and moisture: every optional term of the step."""
    case = Case.bubble3d(20, 16, 14, dx=1000.0, ztop=8000.0, dt=6.0, seed=21, bubble_zc=3000.0)
    g = case.grid
    prm = case.params(time_step_sound=6, mapped=1, coriolis=1, diff_6th_opt=2, diff_6th_factor=0.12,
                      n_moist=1, n_scalars=2, damp_opt=3, zdamp=3000.0)

    def m(xf, yf):
        return 1.0 + 0.05 * np.sin(np.pi * xf / g.nx) * np.cos(np.pi * yf / g.ny)

    ii, jj = np.meshgrid(np.arange(g.nx + 1, dtype=float), np.arange(g.ny + 1, dtype=float))
    mass = m(ii[:g.ny, :g.nx] + 0.5, jj[:g.ny, :g.nx] + 0.5)[None]
    uf = m(ii[:g.ny, :], jj[:g.ny, :] + 0.5)[None]
    vf = m(ii[:, :g.nx] + 0.5, jj[:, :g.nx])[None]
    rng = np.random.default_rng(5)
    shp = A.field_shape(A.MI_Q, g.nx, g.ny, g.nz)
    lat = np.deg2rad(30 + 5 * (jj[:g.ny, :g.nx] / g.ny))[None]
    fields = {
        A.MI_MSFTX: mass, A.MI_MSFTY: 1.01 * mass, A.MI_MSFUX: uf, A.MI_MSFUY: 1.01 * uf, A.MI_MSFVX: vf,
        A.MI_MSFVY: 1.01 * vf, A.MI_MSFVX_INV: 1.0 / vf,
        A.MI_F: 2 * 7.292e-5 * np.sin(lat), A.MI_E: 2 * 7.292e-5 * np.cos(lat),
        A.MI_SINA: 0.1 * np.ones_like(mass), A.MI_COSA: np.sqrt(0.99) * np.ones_like(mass),
        A.MI_Q + 0: 0.01 * np.exp(-np.arange(g.nz) / 5.0)[:, None, None] * (1 + 0.1 * rng.random(shp)),
        A.MI_Q + 1: (rng.random(shp) > 0.6).astype(float),
        A.MI_Q_TENDF + 0: 1e-5 * rng.standard_normal(shp),
    }
    return case, prm, fields


def test_mapped_coriolis_diffusion_phases_bitwise(ctx):
    print("[parity] mapped + coriolis + sixth-order diffusion + moisture phases", flush=True)
    case, prm, fields = _mapped_case()
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    for fid, a in fields.items():
        gpu.set_field(fid, a)
        orc.set_field(fid, a)
    QS = (A.MI_Q, A.MI_Q + 1)
    for step in range(2):
        gpu.step_begin()
        orc.step_begin()
        for rk in (1, 2, 3):
            gpu.rk_prep(rk)
            orc.rk_prep(rk)
            gpu.rk_tendency(rk)
            orc.rk_tendency(rk)
            _compare(gpu, orc, TEND, f"step {step} rk {rk} tendencies")
            gpu.small_steps(rk)
            orc.small_steps(rk)
            _compare(gpu, orc, STAGE, f"step {step} rk {rk} acoustic steps")
            gpu.scalar_transport(rk)
            orc.scalar_transport(rk)
            _compare(gpu, orc, QS, f"step {step} rk {rk} scalar transport")
            gpu.calc_p_rho_phi(rk)
            orc.calc_p_rho_phi(rk)
    gpu.step(5)
    orc.step(5)
    _compare(gpu, orc, STATE + QS, "after 5 graph steps")


def test_hills_bitwise(ctx):
    """The real-data-shaped domain: terrain (full base state, surface w'' and set_w_surface), hybrid
    coordinate, map factors, Coriolis, moisture with the PD limiter, sixth-order diffusion and the
    Rayleigh layer, phase by phase over one step and then 6 graph-replayed steps."""
    print("[parity] hills", flush=True)
    case = Case.hills(24, 20, 16, dx=15000.0, dt=90.0, seed=5)
    prm = case.params(time_step_sound=4, damp_opt=3, zdamp=5000.0, diff_6th_opt=2, n_scalars=2)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    g = case.grid
    tr = np.random.default_rng(8).random(A.field_shape(A.MI_Q, g.nx, g.ny, g.nz))
    gpu.set_field(A.MI_Q + 1, tr)
    orc.set_field(A.MI_Q + 1, tr)
    QS = (A.MI_Q, A.MI_Q + 1)
    gpu.step_begin()
    orc.step_begin()
    for rk in (1, 2, 3):
        for ph in ("rk_prep", "rk_tendency", "small_steps", "scalar_transport", "calc_p_rho_phi"):
            getattr(gpu, ph)(rk)
            getattr(orc, ph)(rk)
        _compare(gpu, orc, STATE + QS + TEND, f"hills stage {rk}")
    # the phase API leaves set_w_surface to the step: run it through mi_step from here on
    gpu2 = RK3(ctx, case.grid, prm, case)
    orc2 = OracleDomain(case, prm)
    for d in (gpu2, orc2):
        d.set_field(A.MI_Q + 1, tr)
    for chunk in (1, 5):
        gpu2.step(chunk)
        orc2.step(chunk)
        _compare(gpu2, orc2, STATE + QS, f"hills after a chunk of {chunk} steps")
    assert np.abs(gpu2.field(A.MI_W)[0]).max() > 0  # the surface w follows the terrain


def test_cfl_max_equals_oracle(ctx):
    """mi_cfl_max between the tendencies and the acoustic steps (the reference's on_cfl point) equals the
    oracle's maxima exactly, on the hills domain."""
    case = Case.hills(30, 24, 20, dx=15000.0, dt=90.0, seed=6)
    prm = case.params(time_step_sound=4)
    gpu = RK3(ctx, case.grid, prm, case)
    orc = OracleDomain(case, prm)
    gpu.step(2)
    orc.step(2)
    gpu.step_begin()
    orc.step_begin()
    for d in (gpu, orc):
        d.rk_prep(1)
        d.rk_tendency(1)
    g, o = gpu.cfl_max(), orc.cfl_max()
    assert g == o and g[1] > 0 and g[0] > 0, (g, o)
