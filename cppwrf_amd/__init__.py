"""cppwrf_amd: WRF-ARW's RK3 split-explicit dynamics step on AMD MI355X (gfx950).

The native library (libcppwrf_amd.so: HIP kernels + the C ABI of include/cppwrf_amd.h) does all the
work; this package only loads it and gives the C ABI Python names.  There is no fallback: without the
built library every entry point raises ImportError.
"""
from ._abi import lib  # noqa: F401

__all__ = ["lib"]
