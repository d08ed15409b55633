"""Builds the native library in-tree: cppwrf_amd/libcppwrf_amd.so (HIP for gfx950 + host C++).

Every compile uses -ffp-contract=off (no fused multiply-add: the FP64 mode reproduces the reference
CPU build bit for bit) and no fast-math.  Device code is also built without packed FP32 instructions
(-packed-fp32-ops).  Run as `python -m cppwrf_amd.build` or through __graft_entry__.build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcppwrf_amd.so")
ARCH = os.environ.get("MI_OFFLOAD_ARCH", "gfx950")

# (source, object, defines): the kernels and the domain are compiled once per arithmetic
PRECISIONS = [("fp64", ["-DMI_REAL=double", "-DMI_NS=mi_fp64"]), ("fp32", ["-DMI_REAL=float", "-DMI_NS=mi_fp32"])]
HIP_UNITS = [(src, f"{src}.{tag}.o", defs) for src in ("kernels_fp64.hip", "domain.hip") for tag, defs in PRECISIONS] + \
            [("capi.hip", "capi.hip.o", [])]
HOST_SOURCES = ["cases.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
          "-Wno-unused-parameter"]
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                      "-munsafe-fp-atomics"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    noise = [ln for ln in (r.stdout + r.stderr).splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "target-feature" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build_library(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src, obj, defs in HIP_UNITS:
        jobs.append([HIPCC, *HIP_FLAGS, *defs, "-c", "-o", os.path.join(BUILD, obj), os.path.join(CSRC, src)])
    for s in HOST_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append(["g++", *COMMON, "-c", "-o", obj, os.path.join(CSRC, s)])
    with ThreadPoolExecutor(max_workers=6) as ex:
        list(ex.map(_run, jobs))
    objs = [os.path.join(BUILD, obj) for _, obj, _ in HIP_UNITS] + [os.path.join(BUILD, s + ".o") for s in HOST_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    path = os.path.join(ROOT, "oracle", "libmi_oracle.so")
    if verbose:
        print("built", path)
    return path


def main() -> None:
    build_library(verbose=True)
    build_oracle(verbose=True)


if __name__ == "__main__":
    main()
