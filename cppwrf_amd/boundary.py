"""Boundary data of specified lateral boundaries (the slabs of mi_bdy_set) built from a domain's state.

A real-data run reads these from its boundary file (WRF wrfbdy: values at the start of each interval and
their tendencies over it).  The synthetic real-data-shaped workload takes the initial state's outer ring,
coupled as a boundary file holds it, as the values, and a seeded slow oscillation (period `period_s`,
default 6 h) of the winds, theta and qv as the tendencies, so the boundaries drift smoothly away from the
initial state.  Slab layout (reference LateralBC, src/dyn/lateral_bc.hpp:17-19): (b, k, pos) with pos
fastest, b = 0 the outermost row; pos = j on the x sides, i on the y sides.
"""
from __future__ import annotations

import numpy as np

from . import _abi as A

BW = A.MI_SPEC_BDY_WIDTH


def ring(field: np.ndarray, side: int) -> np.ndarray:
    """the (b, k, pos) slab of the outer BW rows of a (k, j, i) array on one side"""
    if side == A.MI_BDY_XS:
        s = field[:, :, :BW]                     # (k, j, b)
    elif side == A.MI_BDY_XE:
        s = field[:, :, ::-1][:, :, :BW]         # i = last - b
    elif side == A.MI_BDY_YS:
        s = field[:, :BW, :].transpose(0, 2, 1)  # (k, i, b)
    else:
        s = field[:, ::-1, :][:, :BW, :].transpose(0, 2, 1)
    return np.ascontiguousarray(s.transpose(2, 0, 1))


def coupled_state(case, state: dict, levels: dict) -> dict:
    """the fields as a boundary file couples them, from the uncoupled state of a specified-boundary grid:
    u (c1h muu + c2h) / msfuy, v (c1h muv + c2h) / msfvx, t, q (c1h mut + c2h), ph (c1f mut + c2f), mu'"""
    mu, mub = state[A.MI_MU][0], state[A.MI_MUB][0]
    m = mu + mub                                    # (j, i)
    mu_u = np.empty((m.shape[0], m.shape[1] + 1))  # specified edges: the column taken twice
    mu_u[:, 1:-1] = 0.5 * (m[:, 1:] + m[:, :-1])
    mu_u[:, 0], mu_u[:, -1] = m[:, 0], m[:, -1]
    mu_v = np.empty((m.shape[0] + 1, m.shape[1]))
    mu_v[1:-1, :] = 0.5 * (m[1:, :] + m[:-1, :])
    mu_v[0, :], mu_v[-1, :] = m[0, :], m[-1, :]
    c1h, c2h, c1f, c2f = (levels[k][:, None, None] for k in ("c1h", "c2h", "c1f", "c2f"))
    out = {
        A.MI_BDY_U: state[A.MI_U] * (c1h * mu_u[None] + c2h) / state[A.MI_MSFUY],
        A.MI_BDY_V: state[A.MI_V] * (c1h * mu_v[None] + c2h) / state[A.MI_MSFVX],
        A.MI_BDY_T: state[A.MI_T] * (c1h * m[None] + c2h),
        A.MI_BDY_PH: state[A.MI_PH] * (c1f * m[None] + c2f),
        A.MI_BDY_MU: mu[None],
    }
    for n, q in state.get("q", {}).items():
        out[A.MI_BDY_Q + n] = q * (c1h * m[None] + c2h)
    return out


def level_arrays(case) -> dict:
    lv, nz = case.levels, case.grid.nz
    return {k: np.ctypeslib.as_array(getattr(lv, k), shape=(nz + (1 if k[-1] == "f" else 0),)).copy()
            for k in ("c1h", "c2h", "c1f", "c2f")}


def drifting_boundary(case, seed: int = 7, period_s: float = 6 * 3600.0, amp_wind: float = 2.0, amp_theta: float = 0.5,
                      amp_qv: float = 0.05, q_species=(0,)):
    """{(var, side): (values, tendencies)} for every boundary variable and the scalars in q_species: the
    initial ring as values, and tendencies of a slow oscillation X + a sin(2 pi t / period + phase) at t = 0
    (winds amp_wind m/s, theta amp_theta K, qv amp_qv of itself; ph and mu held), with phases seeded along
    each side"""
    g = case.grid

    def field(fid):
        try:
            return case.field(fid)
        except A.MiError:  # a case without map factors: 1
            return np.ones(A.field_shape(fid, g.nx, g.ny, g.nz))

    st = {fid: field(fid) for fid in (A.MI_U, A.MI_V, A.MI_T, A.MI_PH, A.MI_MU, A.MI_MUB, A.MI_MSFUY, A.MI_MSFVX)}
    st = {k: (v if v.ndim == 3 else v.reshape(1, *v.shape[-2:])) for k, v in st.items()}
    if case.moist:
        st["q"] = {n: case.field(A.MI_Q + n) for n in q_species}
    cpl = coupled_state(case, st, level_arrays(case))
    ones = {k: np.ones_like(v) for k, v in st.items() if k != "q"}
    unit = coupled_state(case, {**ones, A.MI_MU: st[A.MI_MU], A.MI_MUB: st[A.MI_MUB],
                                A.MI_MSFUY: st[A.MI_MSFUY], A.MI_MSFVX: st[A.MI_MSFVX],
                                "q": {n: np.ones_like(q) for n, q in st.get("q", {}).items()}}, level_arrays(case))
    rng = np.random.default_rng(seed)
    w = 2.0 * np.pi / period_s
    amp = {A.MI_BDY_U: amp_wind, A.MI_BDY_V: amp_wind, A.MI_BDY_T: amp_theta, A.MI_BDY_PH: 0.0, A.MI_BDY_MU: 0.0}
    out = {}
    for var, val in cpl.items():
        for side in range(4):
            v = ring(val, side)
            b_, nk, npos = v.shape
            assert (b_, nk, npos) == A.bdy_shape(var if var < A.MI_BDY_Q else A.MI_BDY_T, side, g.nx, g.ny, g.nz)
            phase = np.cos(2.0 * np.pi * np.arange(npos) / npos + rng.uniform(0.0, 2.0 * np.pi))[None, None, :]
            if var >= A.MI_BDY_Q:
                t = amp_qv * w * phase * v  # relative amplitude of the coupled qv itself
            else:
                t = amp[var] * w * phase * ring(unit[var], side)
            out[(var, side)] = (v, np.ascontiguousarray(t))
    return out
