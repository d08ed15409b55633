"""The device ports of glibc pow, exp, sin and cos (cppwrf_amd/csrc/libm_port.hpp) against the host glibc, bit for
bit.

The same header is compiled for the host here (g++, FMA contraction off); the GPU build compiles the identical
expression trees with explicit fma calls at the sites glibc's own build contracts, so a host mismatch would be a
device mismatch too.  The GPU parity tests then cover the device build end to end (pow in calc_p_rho_phi, sin in
the Rayleigh damping layers, cos in the damping of the eddy viscosities, pow and exp in the Kessler
microphysics)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_pow_exp_sin_cos_ports_match_host_glibc(tmp_path):
    exe = tmp_path / "libm_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(HERE, "native", "libm_check.cpp")], check=True)
    r = subprocess.run([str(exe), "3000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
