"""The allocation-free CFL maxima (mi_cfl_max_noalloc, through the library's file-scope scratch array) equal
mi_cfl_max's exactly, called through the C ABI on two calls in a row, the second on a smaller domain."""
import ctypes as C

import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd._abi import check, lib
from cppwrf_amd.dyn import Case, Context, RK3

pytestmark = pytest.mark.gpu


def _both(d):
    v0, h0, v1, h1 = C.c_double(-1), C.c_double(-1), C.c_double(-2), C.c_double(-2)
    check(lib().mi_cfl_max(d.handle, C.byref(v0), C.byref(h0)))
    check(lib().mi_cfl_max_noalloc(d.handle, C.byref(v1), C.byref(h1)))
    return (v0.value, h0.value), (v1.value, h1.value)


@pytest.mark.parametrize("precision", [A.MI_FP64, A.MI_FP32_BULK])
def test_cfl_noalloc_equals_cfl(precision):
    print("[no_alloc] cfl maxima, two domains", flush=True)
    ctx = Context(0)
    big = Case.hills(48, 40, 20, dx=15000.0, dt=90.0, seed=5)
    small = Case.bubble3d(20, 16, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=3, bubble_zc=2500.0)
    out = []
    for case, kw in ((big, dict(time_step_sound=4, damp_opt=3)), (small, dict(time_step_sound=4))):
        d = RK3(ctx, case.grid, case.params(precision=precision, **kw), case)
        d.step(2)
        ref, got = _both(d)
        assert ref == got, (ref, got)
        assert ref[0] > 0 and ref[1] > 0
        out.append(got)
        d.close()
    assert out[0] != out[1]  # the second call's values are its own domain's, not the first's left over
