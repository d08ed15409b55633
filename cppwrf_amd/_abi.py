"""ctypes mirror of include/cppwrf_amd.h and the loader of the native library.

The library is built in-tree (cppwrf_amd/libcppwrf_amd.so, see cppwrf_amd/build.py).  Loading never
falls back to anything else: a missing or stale library is an ImportError naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcppwrf_amd.so")

MI_OK = 0
MI_EINVAL = -1
MI_EHIP = -2
MI_ENOMEM = -3
MI_ENONFINITE = -4
MI_ENODEV = -5

MI_BC_PERIODIC, MI_BC_SYMMETRIC, MI_BC_SPECIFIED, MI_BC_OPEN = 0, 1, 2, 3
MI_BATCH_BRANCHES, MI_BATCH_FUSED = 0, 1
MI_FP64, MI_FP32_BULK = 0, 1

# field ids
MI_U, MI_V, MI_W, MI_T, MI_PH, MI_MU, MI_P, MI_AL = range(8)
MI_WW, MI_RU_M, MI_RV_M, MI_WW_M, MI_ALT, MI_PHP, MI_MUTS = range(8, 15)
MI_RU_TEND, MI_RV_TEND, MI_RW_TEND, MI_T_TEND, MI_PH_TEND = range(15, 20)
MI_MUB, MI_PB, MI_ALB, MI_PHB, MI_T_INIT = range(32, 37)
MI_MSFTX, MI_MSFTY, MI_MSFUX, MI_MSFUY, MI_MSFVX, MI_MSFVY, MI_MSFVX_INV = range(40, 47)
MI_CQU, MI_CQV, MI_CQW, MI_H_DIABATIC = range(20, 24)
MI_RU_TENDF, MI_RV_TENDF, MI_RW_TENDF, MI_T_TENDF, MI_TH_FORCING, MI_QV_FORCING = range(24, 30)
MI_F, MI_E, MI_SINA, MI_COSA, MI_HT = range(47, 52)
MI_U_BASE, MI_V_BASE, MI_QV_BASE, MI_RAINNC, MI_RAINNCV, MI_XKMH, MI_XKHH, MI_T_BASE, MI_Z_BASE = range(52, 61)
MI_Q, MI_Q_TENDF = 64, 96
MI_MAX_MOIST, MI_MAX_SCALARS = 8, 16
ABI_VERSION = 5

MI_CASE_REST_STATE, MI_CASE_WARM_BUBBLE, MI_CASE_IGW, MI_CASE_BUBBLE3D, MI_CASE_HILLS, MI_CASE_STRAKA, MI_CASE_SQUALL2D = range(7)

# specified lateral boundaries
MI_BDY_U, MI_BDY_V, MI_BDY_T, MI_BDY_PH, MI_BDY_MU, MI_BDY_Q = 0, 1, 2, 3, 4, 16
MI_BDY_XS, MI_BDY_XE, MI_BDY_YS, MI_BDY_YE = range(4)
MI_SPEC_BDY_WIDTH = 5


def bdy_shape(var: int, side: int, nx: int, ny: int, nz: int) -> tuple[int, int, int]:
    """(b, k, pos) shape of one side's boundary slab of a variable (pos fastest in memory)."""
    if side in (MI_BDY_XS, MI_BDY_XE):
        npos = ny + (1 if var == MI_BDY_V else 0)
    else:
        npos = nx + (1 if var == MI_BDY_U else 0)
    nk = {MI_BDY_PH: nz + 1, MI_BDY_MU: 1}.get(var, nz)
    return (MI_SPEC_BDY_WIDTH, nk, npos)

# staggering of each field id: (x extra, y extra, z extra, rank)
STAGGER = {
    MI_U: "U", MI_V: "V", MI_W: "W", MI_T: "M", MI_PH: "W", MI_MU: "M2", MI_P: "M", MI_AL: "M",
    MI_WW: "W", MI_RU_M: "U", MI_RV_M: "V", MI_WW_M: "W", MI_ALT: "M", MI_PHP: "M", MI_MUTS: "M2",
    MI_RU_TEND: "U", MI_RV_TEND: "V", MI_RW_TEND: "W", MI_T_TEND: "M", MI_PH_TEND: "W",
    MI_MUB: "M2", MI_PB: "M", MI_ALB: "M", MI_PHB: "W", MI_T_INIT: "M",
    MI_CQU: "U", MI_CQV: "V", MI_CQW: "W", MI_H_DIABATIC: "M",
    MI_RU_TENDF: "U", MI_RV_TENDF: "V", MI_RW_TENDF: "W", MI_T_TENDF: "M", MI_TH_FORCING: "M", MI_QV_FORCING: "M",
    MI_MSFTX: "M2", MI_MSFTY: "M2", MI_MSFUX: "U2", MI_MSFUY: "U2", MI_MSFVX: "V2", MI_MSFVY: "V2",
    MI_MSFVX_INV: "V2", MI_F: "M2", MI_E: "M2", MI_SINA: "M2", MI_COSA: "M2", MI_HT: "M2",
    MI_U_BASE: "Z", MI_V_BASE: "Z", MI_QV_BASE: "Z", MI_RAINNC: "M2", MI_RAINNCV: "M2", MI_XKMH: "M", MI_XKHH: "M", MI_T_BASE: "Z", MI_Z_BASE: "Z",
}
for _n in range(MI_MAX_SCALARS):
    STAGGER[MI_Q + _n] = "M"
    STAGGER[MI_Q_TENDF + _n] = "M"

FIELD_NAMES = {
    MI_U: "u", MI_V: "v", MI_W: "w", MI_T: "t", MI_PH: "ph", MI_MU: "mu", MI_P: "p", MI_AL: "al",
    MI_WW: "ww", MI_RU_M: "ru_m", MI_RV_M: "rv_m", MI_WW_M: "ww_m", MI_ALT: "alt", MI_PHP: "php",
    MI_MUTS: "muts", MI_RU_TEND: "ru_tend", MI_RV_TEND: "rv_tend", MI_RW_TEND: "rw_tend",
    MI_T_TEND: "t_tend", MI_PH_TEND: "ph_tend", MI_CQU: "cqu", MI_CQV: "cqv", MI_CQW: "cqw",
    MI_H_DIABATIC: "h_diabatic", MI_RU_TENDF: "ru_tendf", MI_RV_TENDF: "rv_tendf", MI_RW_TENDF: "rw_tendf",
    MI_T_TENDF: "t_tendf", MI_TH_FORCING: "rthften", MI_QV_FORCING: "rqvften", MI_RAINNC: "rainnc",
    MI_RAINNCV: "rainncv", MI_H_DIABATIC: "h_diabatic", MI_XKMH: "xkmh", MI_XKHH: "xkhh", MI_T_BASE: "t_base", MI_Z_BASE: "z_base",
}
for _n in range(MI_MAX_SCALARS):
    FIELD_NAMES[MI_Q + _n] = f"q{_n}"
    FIELD_NAMES[MI_Q_TENDF + _n] = f"q{_n}_tendf"


def field_shape(fid: int, nx: int, ny: int, nz: int) -> tuple[int, int, int]:
    """(nzs, nys, nxs) of a field's logical points, the dense array layout of the C ABI."""
    st = STAGGER[fid]
    if st == "Z":  # a 1-D profile of the mass levels
        return (nz, 1, 1)
    if st == "M2":
        return (1, ny, nx)
    if st == "U2":
        return (1, ny, nx + 1)
    if st == "V2":
        return (1, ny + (1 if ny > 1 else 0), nx)
    nxs = nx + (1 if st == "U" else 0)
    nys = ny + (1 if st == "V" and ny > 1 else 0)
    nzs = nz + (1 if st == "W" else 0)
    return (nzs, nys, nxs)


class MiGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("bc_x", C.c_int32),
                ("bc_y", C.c_int32), ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double)]


class MiParams(C.Structure):
    _fields_ = [("h_adv_order", C.c_int32), ("v_adv_order", C.c_int32), ("top_lid", C.c_int32),
                ("w_damping", C.c_int32), ("w_crit_cfl", C.c_double), ("epssm", C.c_double),
                ("smdiv", C.c_double), ("emdiv", C.c_double), ("damp_opt", C.c_int32),
                ("hypsometric_opt", C.c_int32), ("zdamp", C.c_double), ("dampcoef", C.c_double),
                ("time_step_sound", C.c_int32), ("precision", C.c_int32), ("n_moist", C.c_int32),
                ("n_scalars", C.c_int32), ("moist_adv_opt", C.c_int32), ("scalar_adv_opt", C.c_int32),
                ("mapped", C.c_int32), ("coriolis", C.c_int32), ("diff_6th_opt", C.c_int32), ("terrain", C.c_int32),
                ("diff_6th_factor", C.c_double), ("diff_opt", C.c_int32), ("km_opt", C.c_int32),
                ("khdif", C.c_double), ("kvdif", C.c_double), ("pr_inv", C.c_double), ("mp_physics", C.c_int32),
                ("mp_tend_lim", C.c_double), ("c_s", C.c_double)]


_PD = C.POINTER(C.c_double)


class MiLevels(C.Structure):
    _fields_ = [(n, _PD) for n in ("dnw", "rdnw", "dn", "rdn", "fnm", "fnp", "c1h", "c2h", "c3h", "c4h",
                                   "c1f", "c2f", "c3f", "c4f")] + \
               [(n, C.c_double) for n in ("cf1", "cf2", "cf3", "cfn", "cfn1", "rdx", "rdy", "p_top")]


class MiCaseSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("grid", MiGrid), ("ztop", C.c_double), ("seed", C.c_uint64),
                ("noise_amp", C.c_double), ("bubble_dtheta", C.c_double), ("bubble_radius", C.c_double),
                ("bubble_zc", C.c_double), ("u0", C.c_double), ("n_hills", C.c_int32), ("moist", C.c_int32),
                ("hill_height", C.c_double), ("hill_wmin", C.c_double), ("hill_wmax", C.c_double),
                ("jet_min", C.c_double), ("jet_max", C.c_double), ("map_amp", C.c_double), ("etac", C.c_double),
                ("p_top", C.c_double), ("qv0", C.c_double), ("qv_hscale", C.c_double)]


def default_params(**kw) -> MiParams:
    """The reference's defaults (src/cases/model.cpp:256-281)."""
    p = MiParams(h_adv_order=5, v_adv_order=3, top_lid=0, w_damping=0, w_crit_cfl=1.0, epssm=0.1, smdiv=0.1,
                 emdiv=0.01, damp_opt=0, hypsometric_opt=1, zdamp=5000.0, dampcoef=0.2, time_step_sound=0,
                 precision=MI_FP64, n_moist=0, n_scalars=0, moist_adv_opt=1, scalar_adv_opt=1, mapped=0,
                 coriolis=0, diff_6th_opt=0, terrain=0, diff_6th_factor=0.12, diff_opt=0, km_opt=1, khdif=0.0,
                 kvdif=0.0, pr_inv=3.0, mp_physics=0, mp_tend_lim=10.0, c_s=0.25)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _proto(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_VP = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
# mi_cfl_fn / mi_stage_fn
CFL_FN = C.CFUNCTYPE(None, _VP, C.c_double, C.c_double)
STAGE_FN = C.CFUNCTYPE(None, _VP)
class MiFieldView(C.Structure):
    """mi_field_view"""
    _fields_ = [("data", _VP), ("elem_bytes", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("stride_k", C.c_int64), ("stride_j", C.c_int64)]


class MiFieldLayout(C.Structure):
    """mi_field_layout"""
    _fields_ = [("elem_bytes", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("stride_k", C.c_int64), ("stride_j", C.c_int64), ("offset", C.c_int64), ("alloc_elems", C.c_int64)]


# mi_phys_fn (user, domain)
PHYS_FN = C.CFUNCTYPE(None, _VP, _VP)


class MiColumnPhysics(C.Structure):
    """mi_column_physics"""
    _fields_ = [("user", _VP), ("part1", PHYS_FN), ("part2", PHYS_FN), ("after_rk", PHYS_FN),
                ("after_all_rk_steps", PHYS_FN), ("th_forcing", C.c_int32), ("qv_forcing", C.c_int32)]


# name -> (restype, argtypes), every function include/cppwrf_amd.h declares
SIGNATURES = {
    "mi_last_error": (C.c_char_p, []),
    "mi_abi_version": (_I, []),
    "mi_auto_sound_steps": (_I, [C.c_double, C.c_double, C.c_double]),
    "mi_adaptive_sound_steps": (_I, [C.c_double, C.c_double, C.c_double, C.c_double]),
    "mi_ctx_create": (_I, [_I, _VP, C.POINTER(_VP)]),
    "mi_ctx_destroy": (_I, [_VP]),
    "mi_domain_create": (_I, [_VP, C.POINTER(MiGrid), C.POINTER(MiParams), C.POINTER(MiLevels), C.POINTER(_VP)]),
    "mi_domain_destroy": (_I, [_VP]),
    "mi_domain_n_sound": (_I, [_VP]),
    "mi_domain_device_bytes": (_I64, [_VP]),
    "mi_domain_set_field": (_I, [_VP, _I, _PD, _I64]),
    "mi_domain_get_field": (_I, [_VP, _I, _PD, _I64]),
    "mi_step": (_I, [_VP, _I]),
    "mi_domain_sync": (_I, [_VP]),
    "mi_step_begin": (_I, [_VP]),
    "mi_rk_prep": (_I, [_VP, _I]),
    "mi_rk_tendency": (_I, [_VP, _I]),
    "mi_small_steps": (_I, [_VP, _I]),
    "mi_scalar_transport": (_I, [_VP, _I]),
    "mi_calc_p_rho_phi": (_I, [_VP, _I]),
    "mi_calc_p_rho_phi_interior": (_I, [_VP, _I]),
    "mi_domain_set_diffusion": (_I, [_VP, _I, C.c_double]),
    "mi_step_end": (_I, [_VP]),
    "mi_bdy_set": (_I, [_VP, _I, _I, _PD, _PD, _I64]),
    "mi_bdy_set_time": (_I, [_VP, C.c_double]),
    "mi_bdy_get_time": (_I, [_VP, C.POINTER(C.c_double)]),
    "mi_batch_create": (_I, [_VP, C.POINTER(_VP), _I, C.POINTER(_VP)]),
    "mi_batch_destroy": (_I, [_VP]),
    "mi_batch_step": (_I, [_VP, _I]),
    "mi_batch_graph_count": (_I, [_VP]),
    "mi_batch_set_schedule": (_I, [_VP, _I]),
    "mi_cfl_max": (_I, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mi_cfl_max_noalloc": (_I, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mi_domain_set_dt": (_I, [_VP, C.c_double, _I]),
    "mi_domain_state_bytes": (_I64, [_VP]),
    "mi_domain_save_state": (_I, [_VP, _VP, _I64]),
    "mi_domain_load_state": (_I, [_VP, _VP, _I64]),
    "mi_domain_set_cfl_callback": (_I, [_VP, CFL_FN, _VP]),
    "mi_domain_set_first_rk_prep_callback": (_I, [_VP, STAGE_FN, _VP]),
    "mi_domain_set_column_physics": (_I, [_VP, C.POINTER(MiColumnPhysics)]),
    "mi_domain_field_view": (_I, [_VP, _I, C.POINTER(MiFieldView)]),
    "mi_domain_field_layout": (_I, [_VP, _I, C.POINTER(MiFieldLayout)]),
    "mi_field_bind": (_I, [_VP, _I, _VP, _I64]),
    "mi_device_alloc": (_I, [_VP, _I64, C.POINTER(_VP)]),
    "mi_device_free": (_I, [_VP, _VP]),
    "mi_device_upload": (_I, [_VP, _VP, _VP, _I64]),
    "mi_device_download": (_I, [_VP, _VP, _VP, _I64]),
    "mi_check_finite": (_I, [_VP, C.POINTER(_I)]),
    "mi_domain_autotune": (_I, [_VP, _I]),
    "mi_domain_get_tuning": (_I, [_VP, C.POINTER(_I), _I]),
    "mi_domain_set_tuning": (_I, [_VP, C.POINTER(_I), _I]),
    "mi_profile_kernel": (_I, [_VP, C.c_char_p]),
    "mi_profile_read": (_I, [_VP, C.POINTER(C.c_double), C.POINTER(_I), C.POINTER(C.c_double)]),
    "mi_case_create": (_I, [C.POINTER(MiCaseSpec), C.POINTER(_VP)]),
    "mi_case_destroy": (_I, [_VP]),
    "mi_case_levels": (_I, [_VP, C.POINTER(MiLevels)]),
    "mi_case_get_field": (_I, [_VP, _I, _PD, _I64]),
    "mi_case_load": (_I, [_VP, _VP]),
}

_lib = None


def lib():
    """The native library (loaded once).  Raises ImportError if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            _proto(L, name, res, args)
        if L.mi_abi_version() != ABI_VERSION:
            raise ImportError("libcppwrf_amd.so ABI version mismatch: rebuild it")
        _lib = L
    return _lib


class MiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != MI_OK:
        raise MiError(rc, lib().mi_last_error().decode())
