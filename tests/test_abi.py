"""The C-ABI library and its host-side planner (CPU only: no compute calls).

- libvsip_amd.so loads and exports every function include/vsip.h and
  include/vsip_amd.h declare;
- the planner (vsip_amd_schedule, pure host code) gives every supported
  length a schedule whose launch shapes respect the kernel's register and LDS
  rules, and rejects lengths with prime factors above 13;
- without a device, constructors fail cleanly (NULL / error code), as the
  reference's do on allocation failure.
"""
import ctypes
import os
import re
import subprocess

import pytest

from jvsip_amd.vsip import LIB_PATH, AmdLib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vsip_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    os.environ.setdefault("VSIP_AMD_QUIET", "1")
    return AmdLib(LIB_PATH)


def test_exports_every_declared_symbol(lib):
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB_PATH]).decode()
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    declared = _declared("vsip.h") + _declared("vsip_amd.h")
    assert len(declared) > 100
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        getattr(lib.dll, s)


def test_header_is_c89():
    src = "#include \"vsip.h\"\n#include \"vsip_amd.h\"\nint main(void){return 0;}\n"
    r = subprocess.run(["gcc", "-std=c89", "-pedantic", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        "-x", "c", "-", "-o", os.devnull], input=src.encode(), capture_output=True)
    assert r.returncode == 0, r.stderr.decode()


def _kernel_rules_hold(s, prec, lds_max):
    maxe_ok = {4: (16, 32), 8: (8,)}[prec]
    n = s["n"]
    prod = 1
    for r in s["radix"]:
        assert r in (1, 2, 3, 4, 5, 7, 8, 11, 13) + ((6, 9, 10, 12, 14, 15, 16) if prec == 4 else ())
        prod *= r
    assert prod == n
    S = s["lds_stride"]
    assert S >= n + (n >> (5 if prec == 4 else 4)) and S % 2 == 1
    for vpb, thr, maxe in (s["along"], s["across"]):
        assert maxe in maxe_ok and 64 <= thr <= (512 if maxe > 16 else 1024) and thr % 64 == 0 and vpb >= 1
        assert vpb * S * 2 * prec + vpb * 20 <= lds_max
        for r in s["radix"]:
            per_thread = -(-(vpb * (n // r)) // thr)
            cap = 1 if (r == 6 or r >= 9) and r != 16 else max(1, maxe // r)
            assert per_thread <= cap, (n, r, vpb, thr, maxe)


@pytest.mark.parametrize("prec", [4, 8])
def test_planner_all_small_lengths(lib, prec):
    lds_max = 163840
    planned = 0
    for n in range(1, 3000):
        s = lib.schedule(prec, n, 1000, lds_max)
        m = n
        for p in (2, 3, 5, 7, 11, 13):
            while m % p == 0:
                m //= p
        if m != 1:
            assert s["rc"] == -1, n
            continue
        assert s["rc"] == 0, (n, s)
        _kernel_rules_hold(s, prec, lds_max)
        planned += 1
    assert planned > 400


@pytest.mark.parametrize("n,prec", [(4096, 4), (16384, 4), (8192, 4), (19683, 4), (10000, 4),
                                    (8192, 8), (6561, 8), (65536, 4), (40000, 4), (16384, 8)])
def test_planner_large_lengths(lib, n, prec):
    s = lib.schedule(prec, n, 4096, 163840)
    if s["rc"] == -2:
        assert n * 2 * prec > 163840 // 2  # only lengths that cannot fit one workgroup split
        return
    assert s["rc"] == 0
    _kernel_rules_hold(s, prec, 163840)


def test_headline_schedule(lib):
    s = lib.schedule(4, 4096, 4096)
    assert s["radix"] == [16, 16, 16]
    assert s["along"] == [1, 256, 16]


def test_no_device_constructors_fail_cleanly(lib):
    import torch  # noqa: F401  (only to skip where a device is present)
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    assert not lib.ccfftmip_create_f(8, 4, 1.0, -1, 0, 0, 0)
    assert not lib.cblockcreate_f(16, 0)
    assert lib.last_error()
