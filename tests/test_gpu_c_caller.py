"""A plain C89 VSIPL program (tests/c_caller/fftm_caller.c) built against
include/vsip.h and linked against libvsip_amd.so, the way a jvsip user's
program relinks: it must compile cleanly on any machine, and on a device its
ccfftmip / ccfftmop / rcfftmop+crfftmop results must match its own direct
DFT (rel L2 < 1e-5 per case)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_caller", "fftm_caller.c")
LIBDIR = os.path.join(ROOT, "jvsip_amd", "lib")


def _build(tmp_path):
    exe = str(tmp_path / "fftm_caller")
    cmd = ["gcc", "-std=c89", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           SRC, "-L", LIBDIR, "-lvsip_amd", "-Wl,-rpath," + LIBDIR, "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_builds(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libvsip_amd.so")):
        pytest.skip("library not built")
    _build(tmp_path)


@pytest.mark.gpu
def test_c_caller_runs(tmp_path):
    exe = _build(tmp_path)
    print("\n[case] C89 caller: ccfftmip ROW, ccfftmop COL, rc/cr round trip", flush=True)
    sys.stdout.flush()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "ALL OK" in r.stdout
