"""The C++20 facade driven as a reference caller drives dyn::RK3 (tests/native/facade_gpu.cpp): caller-owned
State / Tendencies / BaseState device buffers stepped by RK3::step(State&, Tendencies&, const BaseState&), with a
LateralBC object for specified boundaries; the fields read back from the caller's buffers are bit for bit the
oracle's."""
import os
import subprocess

import numpy as np
import pytest

from cppwrf_amd import _abi as A
from cppwrf_amd.boundary import drifting_boundary
from cppwrf_amd.dyn import Case
from oracle_lib import OracleDomain

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "cppwrf_amd")
STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("facade") / "facade_gpu"
    subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-o", str(out),
                    os.path.join(HERE, "native", "facade_gpu.cpp"), f"-L{LIBDIR}", "-lcppwrf_amd",
                    f"-Wl,-rpath,{LIBDIR}"], check=True)
    return str(out)


def _cases():
    return {
        "bubble": (lambda: Case.bubble3d(20, 16, 12, dx=1000.0, ztop=8000.0, dt=6.0, seed=11, bubble_zc=2500.0),
                   dict(time_step_sound=4, damp_opt=3, zdamp=3000.0), False),
        "hills_specified": (lambda: Case.hills(24, 20, 14, dx=15000.0, dt=90.0, seed=3, bc=A.MI_BC_SPECIFIED),
                            dict(time_step_sound=4, damp_opt=3, diff_6th_opt=1), True),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_facade_steps_caller_buffers_bitwise(exe, tmp_path, name):
    make, kw, spec = _cases()[name]
    print(f"[facade] {name}", flush=True)
    case = make()
    prm = case.params(**kw)
    (tmp_path / "spec.bin").write_bytes(bytes(case.spec))
    (tmp_path / "params.bin").write_bytes(bytes(prm))
    orc = OracleDomain(case, prm)
    if spec:
        bdy = drifting_boundary(case, seed=13)
        with open(tmp_path / "bdy.bin", "wb") as f:
            for (var, side), (v, t) in bdy.items():
                v = np.ascontiguousarray(v, dtype=np.float64).ravel()
                t = np.ascontiguousarray(t, dtype=np.float64).ravel()
                f.write(np.array([var, side], dtype=np.int32).tobytes() + np.array([v.size], dtype=np.int64).tobytes())
                f.write(v.tobytes() + t.tobytes())
        orc.load_boundary(bdy)
    steps = 3
    r = subprocess.run([exe, str(tmp_path), str(steps)], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    orc.step(steps)
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    g = case.grid
    o = 0
    for fid in STATE:
        shp = A.field_shape(fid, g.nx, g.ny, g.nz)
        n = int(np.prod(shp))
        got = raw[o:o + n].reshape(shp)
        o += n
        want = orc.field(fid)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), A.FIELD_NAMES[fid]
    assert o == raw.size
