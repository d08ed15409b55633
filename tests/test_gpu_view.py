"""mi_domain_field_view (the reference's Field3::view, src/core/field.hpp:16-44): the device address and strides it
gives reach the same values mi_domain_get_field copies out, in FP64 and in FP32-bulk storage, for 3-D, staggered,
2-D and scalar fields; a write through the view is what the next get_field reads; fields without a view are
refused."""
import ctypes as C

import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.dyn import Case, Context, RK3

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


D2H, H2D = 2, 1


def _span(v):
    return (v.ny - 1) * v.stride_j + (v.nz - 1) * v.stride_k + v.nx


def read_view(v):
    dt = np.float64 if v.elem_bytes == 8 else np.float32
    buf = np.empty(_span(v), dtype=dt)
    assert hip().hipMemcpy(buf.ctypes.data, v.data, buf.nbytes, D2H) == 0
    k, j, i = np.meshgrid(np.arange(v.nz), np.arange(v.ny), np.arange(v.nx), indexing="ij")
    return buf[i + k * v.stride_k + j * v.stride_j].astype(np.float64)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


@pytest.mark.parametrize("precision", [A.MI_FP64, A.MI_FP32_BULK])
def test_view_reads_the_fields(ctx, precision):
    case = Case.hills(24, 20, 16, dx=15000.0, dt=90.0, seed=3)
    d = RK3(ctx, case.grid, case.params(precision=precision, time_step_sound=4, n_moist=2, n_scalars=2), case)
    d.step(2)
    d.sync()
    for fid in (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_Q + 1, A.MI_MSFUY):
        v = d.field_view(fid)
        assert v.elem_bytes == (8 if precision == A.MI_FP64 or fid in (A.MI_MU, A.MI_MSFUY) else 4)
        got = read_view(v)
        ref = d.field(fid)
        assert got.shape == ref.shape, A.FIELD_NAMES.get(fid, fid)
        assert np.array_equal(got, ref), A.FIELD_NAMES.get(fid, fid)


def test_view_write_is_seen(ctx):
    case = Case.bubble3d(20, 16, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=2)
    d = RK3(ctx, case.grid, case.params(time_step_sound=4), case)
    d.sync()
    v = d.field_view(A.MI_T)
    buf = np.empty(_span(v))
    assert hip().hipMemcpy(buf.ctypes.data, v.data, buf.nbytes, D2H) == 0
    k, j, i = np.meshgrid(np.arange(v.nz), np.arange(v.ny), np.arange(v.nx), indexing="ij")
    buf[i + k * v.stride_k + j * v.stride_j] += 0.25
    before = d.field(A.MI_T)
    assert hip().hipMemcpy(v.data, buf.ctypes.data, buf.nbytes, H2D) == 0
    assert np.array_equal(d.field(A.MI_T), before + 0.25)
    with pytest.raises(A.MiError):
        d.field_view(A.MI_PB)  # a flat base state keeps profiles only
