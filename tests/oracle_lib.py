"""ctypes access to the CPU oracle (oracle/libmi_oracle.so) for the tests and the CPU baseline.

The oracle is test infrastructure: the library under test never calls it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from cppwrf_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "libmi_oracle.so")

_o = None


def olib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_PATH)
        VP, I, PD = C.c_void_p, C.c_int, C.POINTER(C.c_double)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_domain_create": (I, [C.POINTER(A.MiGrid), C.POINTER(A.MiParams), C.POINTER(A.MiLevels), C.POINTER(VP)]),
            "orc_domain_destroy": (I, [VP]),
            "orc_domain_n_sound": (I, [VP]),
            "orc_set_field": (I, [VP, I, PD, C.c_int64]),
            "orc_get_field": (I, [VP, I, PD, C.c_int64]),
            "orc_set_libmvec": (I, [VP, I]),
            "orc_set_threads": (I, [I]),
            "orc_step": (I, [VP, I]),
            "orc_step_begin": (I, [VP]),
            "orc_rk_prep": (I, [VP, I]),
            "orc_rk_tendency": (I, [VP, I]),
            "orc_small_steps": (I, [VP, I]),
            "orc_scalar_transport": (I, [VP, I]),
            "orc_cfl_max": (I, [VP, PD, PD]),
            "orc_set_dt": (I, [VP, C.c_double, I]),
            "orc_calc_p_rho_phi": (I, [VP, I]),
            "orc_calc_p_rho_phi_interior": (I, [VP, I]),
            "orc_flux_eval": (I, [I, PD, C.c_double, C.c_double, PD]),
            "orc_set_diffusion": (I, [VP, I, C.c_double]),
            "orc_step_end": (I, [VP]),
            "orc_set_column_physics": (I, [VP, C.POINTER(OrcColumnPhysics)]),
            "orc_bdy_set": (I, [VP, I, I, PD, PD, C.c_int64]),
            "orc_bdy_set_time": (I, [VP, C.c_double]),
            "orc_bdy_get_time": (I, [VP, PD]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _o = L
    return _o


ORC_PHYS_FN = C.CFUNCTYPE(None, C.c_void_p)


class OrcColumnPhysics(C.Structure):
    _fields_ = [("user", C.c_void_p), ("part1", ORC_PHYS_FN), ("part2", ORC_PHYS_FN), ("after_rk", ORC_PHYS_FN),
                ("after_all_rk_steps", ORC_PHYS_FN), ("th_forcing", C.c_int32), ("qv_forcing", C.c_int32)]


def _ok(rc):
    if rc != 0:
        raise RuntimeError(olib().orc_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleDomain:
    """The oracle's State/BaseState for one case, stepped on the CPU."""

    def __init__(self, case, params, libmvec: bool = False):
        self.grid = case.grid
        h = C.c_void_p()
        _ok(olib().orc_domain_create(C.byref(case.grid.c()), C.byref(params), C.byref(case.levels), C.byref(h)))
        self._h = h
        g = case.grid
        for fid in case.STATE:
            a = case.field(fid)
            _ok(olib().orc_set_field(h, fid, _ptr(a), a.size))
        if getattr(case, "full_base", False):  # terrain: base fields in full, and the real-data fields
            for fid in (A.MI_MUB, A.MI_PB, A.MI_ALB, A.MI_PHB, A.MI_T_INIT) + case.HILLS_EXTRA:
                a = case.field(fid)
                _ok(olib().orc_set_field(h, fid, _ptr(a), a.size))
            if case.moist:
                a = case.field(A.MI_Q)
                _ok(olib().orc_set_field(h, A.MI_Q, _ptr(a), a.size))
        else:
            for fid in (A.MI_PB, A.MI_ALB, A.MI_PHB, A.MI_T_INIT):
                prof = case.profile(fid)
                _ok(olib().orc_set_field(h, fid, _ptr(prof), prof.size))
            mub = case.profile(A.MI_MUB)
            _ok(olib().orc_set_field(h, A.MI_MUB, _ptr(mub), 1))
            if case.moist:
                a = case.field(A.MI_Q)
                _ok(olib().orc_set_field(h, A.MI_Q, _ptr(a), a.size))
        for fid in (A.MI_U_BASE, A.MI_V_BASE, A.MI_QV_BASE, A.MI_T_BASE, A.MI_Z_BASE):  # the 1-D base profiles
            prof = case.field(fid).reshape(-1)
            _ok(olib().orc_set_field(h, fid, _ptr(prof), prof.size))
        _ok(olib().orc_set_libmvec(h, 1 if libmvec else 0))
        self.nx, self.ny, self.nz = g.nx, g.ny, g.nz

    def __del__(self):
        if getattr(self, "_h", None):
            olib().orc_domain_destroy(self._h)
            self._h = None

    @property
    def n_sound(self):
        return olib().orc_domain_n_sound(self._h)

    def field(self, fid):
        a = np.zeros(A.field_shape(fid, self.nx, self.ny, self.nz))
        _ok(olib().orc_get_field(self._h, fid, _ptr(a), a.size))
        return a

    def set_field(self, fid, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _ok(olib().orc_set_field(self._h, fid, _ptr(a), a.size))

    def step(self, n=1):
        _ok(olib().orc_step(self._h, n))

    def step_begin(self):
        _ok(olib().orc_step_begin(self._h))

    def rk_prep(self, rk):
        _ok(olib().orc_rk_prep(self._h, rk))

    def rk_tendency(self, rk):
        _ok(olib().orc_rk_tendency(self._h, rk))

    def small_steps(self, rk):
        _ok(olib().orc_small_steps(self._h, rk))

    def set_dt(self, dt, n_sound):
        _ok(olib().orc_set_dt(self._h, dt, n_sound))

    def cfl_max(self):
        v, h = C.c_double(), C.c_double()
        _ok(olib().orc_cfl_max(self._h, C.byref(v), C.byref(h)))
        return v.value, h.value

    def scalar_transport(self, rk):
        _ok(olib().orc_scalar_transport(self._h, rk))

    def calc_p_rho_phi(self, rk):
        _ok(olib().orc_calc_p_rho_phi(self._h, rk))

    def set_diffusion(self, opt, factor):
        _ok(olib().orc_set_diffusion(self._h, opt, factor))

    def calc_p_rho_phi_interior(self, spec_zone):
        _ok(olib().orc_calc_p_rho_phi_interior(self._h, spec_zone))

    def step_end(self):
        _ok(olib().orc_step_end(self._h))

    def set_column_physics(self, cp):
        """as RK3.set_column_physics: cp's methods get this domain"""
        if cp is None:
            self._phys = None
            _ok(olib().orc_set_column_physics(self._h, None))
            return

        def thunk(name):
            f = getattr(cp, name, None)
            return ORC_PHYS_FN(lambda _u: f(self)) if f is not None else ORC_PHYS_FN()
        self._phys = OrcColumnPhysics(None, thunk("part1"), thunk("part2"), thunk("after_rk"),
                                      thunk("after_all_rk_steps"), 1 if getattr(cp, "th_forcing", False) else 0,
                                      1 if getattr(cp, "qv_forcing", False) else 0)
        _ok(olib().orc_set_column_physics(self._h, C.byref(self._phys)))

    def bdy_set(self, var, side, values, tends):
        v = np.ascontiguousarray(values, dtype=np.float64)
        t = np.ascontiguousarray(tends, dtype=np.float64)
        _ok(olib().orc_bdy_set(self._h, var, side, _ptr(v), _ptr(t), v.size))

    def load_boundary(self, bdy):
        """every slab of a {(var, side): (values, tendencies)} dict"""
        for (var, side), (v, t) in bdy.items():
            self.bdy_set(var, side, v, t)

    @property
    def bdy_time(self):
        d = C.c_double()
        _ok(olib().orc_bdy_get_time(self._h, C.byref(d)))
        return d.value

    @bdy_time.setter
    def bdy_time(self, v):
        _ok(olib().orc_bdy_set_time(self._h, float(v)))
