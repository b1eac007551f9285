"""CPU checks of the kernels compiled at plan creation (no device needed).

The library carries the device headers of the schedule-specialised Stockham
kernel (jvsip_amd/csrc/hip/fftm_static.hpp) and compiles one instance per
plan schedule with hipRTC.  Here the embedded sources must compile for gfx950
for the schedules the planner produces, and schedules without a compiled
form must be refused rather than miscompiled.
"""
import ctypes
import os

import pytest


@pytest.fixture(scope="module")
def lib():
    from jvsip_amd.vsip import AmdLib, LIB_PATH
    os.environ.setdefault("VSIP_AMD_QUIET", "1")
    return AmdLib(LIB_PATH)


@pytest.fixture(scope="module")
def so():
    from jvsip_amd.vsip import LIB_PATH
    d = ctypes.CDLL(LIB_PATH)
    d.vhip_jit_compile_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.POINTER(ctypes.c_size_t)]
    d.vhip_jit_compile_check.restype = ctypes.c_int
    d.vhip_jit_last_error.restype = ctypes.c_char_p
    return d


@pytest.mark.parametrize("prec,n,nvec", [(4, 1000, 4096), (4, 2187, 1024), (4, 1001, 512), (4, 1920, 1000),
                                         (4, 3000, 512), (4, 6561, 64), (4, 100, 10000), (4, 7 * 11 * 13 * 3, 300),
                                         (8, 2048, 2048), (8, 1000, 512), (8, 4096, 64)])
def test_planner_schedules_compile(lib, so, prec, n, nvec):
    s = lib.schedule(prec, n, nvec, 163840)
    assert s["rc"] == 0
    rad = (ctypes.c_int * len(s["radix"]))(*s["radix"])
    for shape, lin, lout in (("along", 0, 0), ("across", 1, 1), ("across", 0, 1), ("along", 2, 2),
                             ("across", 3, 3), ("along", 4, 0), ("across", 5, 1), ("across", 6, 1),
                             ("across", 1, 9), ("along", 16, 0), ("across", 17, 1), ("along", 32, 16),
                             ("across", 33, 17), ("across", 33, 9), ("along", 0, 16)):
        vpb, thr, _ = s[shape]
        size = ctypes.c_size_t()
        rc = so.vhip_jit_compile_check(prec, n, len(s["radix"]), rad, vpb, thr, lin, lout, b"gfx950",
                                       ctypes.byref(size))
        assert rc == 0, so.vhip_jit_last_error()
        assert size.value > 1000


def test_refuses_schedules_without_a_compiled_form(so):
    rad = (ctypes.c_int * 9)(*([2] * 9))
    assert so.vhip_jit_compile_check(4, 512, 9, rad, 1, 64, 0, 0, b"gfx950", None) != 0
    rad = (ctypes.c_int * 2)(4, 4)
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 8, 0, b"gfx950", None) != 0
    assert so.vhip_jit_compile_check(2, 16, 2, rad, 1, 64, 0, 0, b"gfx950", None) != 0
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 4, 2, b"gfx950", None) != 0   # real input + accumulate
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 18, 0, b"gfx950", None) != 0  # complex to real + taper
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 16, 9, b"gfx950", None) != 0  # complex to real + post-twiddle
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 20, 0, b"gfx950", None) != 0  # complex to real + real input
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 36, 0, b"gfx950", None) != 0  # split input + real input
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 34, 0, b"gfx950", None) != 0  # split input + taper
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 1, 24, b"gfx950", None) != 0  # split output + post-twiddle
    assert so.vhip_jit_compile_check(4, 16, 2, rad, 1, 64, 0, 18, b"gfx950", None) != 0  # split output + accumulate
