"""The C ABI: the library loads, exports every function include/cppwrf_amd.h declares, and its host-only
entry points behave (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cppwrf_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mi_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_function():
    L = A.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"libcppwrf_amd.so does not export {n}"
    assert set(names) == set(A.SIGNATURES), "the ctypes table and the header differ"


def test_abi_version():
    assert A.lib().mi_abi_version() == A.ABI_VERSION == 5


@pytest.mark.parametrize("dt,dx,dy,want", [(6.0, 1000.0, 1000.0, 4), (1.0, 100.0, 100.0, 6), (10.0, 2000.0, 2000.0, 4),
                                           (90.0, 15000.0, 15000.0, 4), (18.0, 3000.0, 3000.0, 4), (60.0, 3000.0, 3000.0, 12)])
def test_auto_sound_steps(dt, dx, dy, want):
    # reference RK3::auto_sound_steps (src/dyn/rk3.cpp:69-72): max(2*(int(300 dt/min(dx,dy) - 0.01)+1), 4)
    assert A.lib().mi_auto_sound_steps(dt, dx, dy) == want


@pytest.mark.parametrize("dt,dx,dy,msft,want", [
    (60.0, 15000.0, 15000.0, 1.0, 4),     # 300*60/15000 = 1.2 -> 2*(1+1) = 4
    (60.0, 15000.0, 15000.0, 1.3, 4),     # spacing 11538: int(1.56 - 0.01) = 1 -> 4
    (90.0, 15000.0, 12000.0, 1.05, 6),    # 300*90*1.05/12000 = 2.3625 -> 2*(2+1) = 6
    (10.0, 1000.0, 1000.0, 1.0, 6),       # 3.0 - 0.01 -> 2 -> 6
    (10.0, 1000.0, 1000.0, 1.2, 8),       # 3.6 -> 3 -> 8
])
def test_adaptive_sound_steps(dt, dx, dy, msft, want):
    # reference RK3::adaptive_sound_steps (src/dyn/rk3.cpp:74-77): the spacing divided by the largest map factor
    n = int(300.0 * dt / (min(dx, dy) / msft) - 0.01)
    assert A.lib().mi_adaptive_sound_steps(dt, dx, dy, msft) == max(2 * (n + 1), 4) == want


def test_case_levels_are_consistent():
    c = Case.bubble3d(16, 12, 20, dx=1000.0, ztop=10000.0, dt=6.0)
    lv = c.levels
    nz = 20
    dnw = np.ctypeslib.as_array(lv.dnw, shape=(nz,))
    rdnw = np.ctypeslib.as_array(lv.rdnw, shape=(nz,))
    assert np.all(dnw < 0) and np.allclose(dnw * rdnw, 1.0, rtol=1e-15)
    assert abs(dnw.sum() + 1.0) < 1e-14  # eta goes from 1 to 0
    phb = c.profile(A.MI_PHB)
    assert phb[0] == 0.0 and np.all(np.diff(phb) > 0)
    assert abs(phb[-1] / 9.81 - 10000.0) < 5.0  # model top at ztop


def test_case_noise_is_seeded():
    a = Case.bubble3d(8, 8, 6, dx=1000.0, ztop=6000.0, dt=6.0, seed=5).field(A.MI_T)
    b = Case.bubble3d(8, 8, 6, dx=1000.0, ztop=6000.0, dt=6.0, seed=5).field(A.MI_T)
    c = Case.bubble3d(8, 8, 6, dx=1000.0, ztop=6000.0, dt=6.0, seed=6).field(A.MI_T)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_errors_are_reported():
    L = A.lib()
    h = C.c_void_p()
    bad = A.MiCaseSpec(kind=99, grid=A.MiGrid(nx=4, ny=4, nz=4, dx=1.0, dy=1.0, dt=1.0), ztop=1000.0)
    assert L.mi_case_create(C.byref(bad), C.byref(h)) == A.MI_EINVAL
    assert b"unknown kind" in L.mi_last_error()
    assert L.mi_case_create(None, C.byref(h)) == A.MI_EINVAL


def test_header_constants_and_params_match_python():
    """every `MI_NAME = value` of the header that the Python mirror names has the same value, and mi_params'
    fields come in the header's order with its types (ctypes lays the struct out as the C compiler does)"""
    src = open(os.path.join(ROOT, "include", "cppwrf_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    consts = dict((n, int(v)) for n, v in re.findall(r"\b(MI_[A-Z0-9_]+)\s*=\s*(-?\d+)", src))
    shared = [n for n in consts if hasattr(A, n)]
    assert len(shared) >= 40
    for n in shared:
        assert getattr(A, n) == consts[n], n
    body = re.search(r"typedef struct mi_params \{(.*?)\} mi_params;", src, flags=re.S).group(1)
    fields = []
    for typ, names in re.findall(r"(int32_t|double)\s+([^;]+);", body):
        for nm in names.split(","):
            fields.append((nm.strip(), typ))
    py = [(n, "int32_t" if t is C.c_int32 else "double") for n, t in A.MiParams._fields_]
    assert fields == py
