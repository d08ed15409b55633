"""The C++20 facade with the reference's dyn::RK3 names compiles against the public header, links the library
and reports errors as std::runtime_error (no GPU needed: the no-device path is the one exercised here)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "cppwrf_amd")


def test_facade_compiles_links_and_reports_errors(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libcppwrf_amd.so")):
        pytest.fail("libcppwrf_amd.so is not built: run __graft_entry__.build()")
    exe = tmp_path / "facade_check"
    subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-o", str(exe),
                    os.path.join(HERE, "native", "facade_check.cpp"), f"-L{LIBDIR}", "-lcppwrf_amd",
                    f"-Wl,-rpath,{LIBDIR}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
