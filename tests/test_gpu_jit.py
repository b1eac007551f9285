"""GPU parity of the kernels compiled at plan creation (schedule-specialised
Stockham, fftm_static.hpp), through the C ABI.

With the size threshold at zero every eligible plan (FP32 complex rows, one
pass, not a power of two of the power-of-two kernel) compiles its schedule;
the plan must report the compiled kernel and its results must match the
oracle (rel L2 <= 1e-5) and the schedule-generic kernel (<= 1e-6; the two run
the same arithmetic).
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _check(name, got, want, tol):
    err = po.rel_l2(got, want)
    print("  %s: rel L2 %.3e" % (name, err), flush=True)
    assert err < tol, "%s: rel L2 %.3e >= %.1e" % (name, err, tol)


@pytest.mark.parametrize("n,m", [(3, 7), (12, 5), (100, 33), (143, 9), (243, 10), (1000, 37), (1001, 6),
                                 (1920, 5), (2187, 4), (3125, 3), (6561, 2), (1152, 11), (4000, 3), (60, 200)])
def test_compiled_schedule_matches_oracle_and_generic(amd, monkeypatch, n, m):
    x = datagen.complex_data(95, (m, n))
    v = fftm.dense_view(m, n)
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, VSIP_ROW, 0, 1)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    print("  schedule", d["passes"][0]["radix"], d["passes"][0]["along"], "compiled", d.get("compiled_schedule"),
          flush=True)
    assert d.get("compiled_schedule") is True
    for direction, scale in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / n)):
        _case("ccfftmip_f ROW %dx%d dir %+d (compiled schedule)" % (m, n, direction))
        y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=scale)
        _check("vs oracle", y, po.ccfftm(x.ravel(), v, direction=direction, scale=scale), 1e-5)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.ccfftm(amd, x.ravel(), v)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("vs generic kernel", fftm.ccfftm(amd, x.ravel(), v), g, 1e-6)


def test_compiled_schedule_strided_out_of_place(amd, monkeypatch):
    m, n = 50, 2187
    _case("ccfftmop_f ROW %dx%d, parent rows 2x (compiled schedule)" % (m, n))
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    x = datagen.complex_data(96, (m * 2 * n,))
    xv = fftm.View(0, 2 * n, m, 1, n)
    yv = fftm.View(n, 2 * n, m, 1, n)
    y0 = np.zeros(m * 2 * n, np.complex64)
    xr, y = fftm.ccfftm(amd, x, xv, out_blk=y0, out_view=yv, direction=VSIP_FFT_INV, scale=0.5)
    np.testing.assert_array_equal(xr, x)
    _check("vs oracle", y, po.ccfftm(x, xv, out_blk=y0, out_view=yv, direction=VSIP_FFT_INV, scale=0.5), 1e-5)


def _plan_info(amd, m, n, major):
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, major, 0, 1)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    print("  passes", [(q["n"], q["radix"]) for q in d["passes"]], "compiled", d.get("compiled_schedule"),
          flush=True)
    return d


@pytest.mark.parametrize("m,n", [(1000, 37), (2187, 20), (1920, 300), (1001, 3), (5, 64)])
def test_compiled_schedule_columns(amd, monkeypatch, m, n):
    """COL plans: lanes walk across the columns at the global load and
    store (unit vector stride), along the vector in LDS."""
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    x = datagen.complex_data(97, (m, n))
    v = fftm.dense_view(m, n)
    d = _plan_info(amd, m, n, VSIP_COL)
    assert d.get("compiled_schedule") is True
    for direction, scale in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / m)):
        _case("ccfftmip_f COL %dx%d dir %+d (compiled schedule)" % (m, n, direction))
        y = fftm.ccfftm(amd, x.ravel(), v, major=VSIP_COL, direction=direction, scale=scale)
        _check("vs oracle", y, po.ccfftm(x.ravel(), v, major=VSIP_COL, direction=direction, scale=scale), 1e-5)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.ccfftm(amd, x.ravel(), v, major=VSIP_COL)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("vs generic kernel", fftm.ccfftm(amd, x.ravel(), v, major=VSIP_COL), g, 1e-6)


def test_compiled_schedule_columns_strided_out_of_place(amd, monkeypatch):
    """COL ccfftmop from a column-strided parent into a row-reversed-free
    strided output (element stride 2 vectors, vector stride 1)."""
    m, n = 1000, 40
    _case("ccfftmop_f COL %dx%d, in every other column of a %dx%d parent (compiled schedule)" % (m, n, m, 2 * n))
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    x = datagen.complex_data(98, (m * 2 * n,))
    xv = fftm.View(1, 2 * n, m, 2, n)
    yv = fftm.View(0, n, m, 1, n)
    y0 = np.zeros(m * n, np.complex64)
    xr, y = fftm.ccfftm(amd, x, xv, major=VSIP_COL, out_blk=y0, out_view=yv, direction=VSIP_FFT_INV, scale=0.25)
    np.testing.assert_array_equal(xr, x)
    want = po.ccfftm(x, xv, major=VSIP_COL, out_blk=y0, out_view=yv, direction=VSIP_FFT_INV, scale=0.25)
    _check("vs oracle", y, want, 1e-5)


@pytest.mark.parametrize("m,n", [(3, 59049), (2, 100000), (4, 30000)])
def test_compiled_schedule_two_pass(amd, monkeypatch, m, n):
    """The two-pass split of long rows: pass A (across, post-twiddled) and
    pass B (scratch rows in, across out) both run compiled kernels."""
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    x = datagen.complex_data(99, (m, n))
    v = fftm.dense_view(m, n)
    d = _plan_info(amd, m, n, VSIP_ROW)
    assert len(d["passes"]) == 2 and d.get("compiled_schedule") is True
    x64 = x.astype(np.complex128)
    for direction, scale in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / n)):
        _case("ccfftmip_f ROW %dx%d dir %+d two-pass (compiled schedule)" % (m, n, direction))
        y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=scale).reshape(m, n)
        want = (np.fft.fft(x64, axis=1) if direction < 0 else np.fft.ifft(x64, axis=1) * n) * scale
        _check("vs numpy fp64", y, want, 1e-5)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.ccfftm(amd, x.ravel(), v)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("vs generic kernel", fftm.ccfftm(amd, x.ravel(), v), g, 1e-6)


def test_compiled_schedule_long_1d(amd, monkeypatch):
    n = 1 << 20
    _case("ccfftip_f n=2^20 (two-pass, compiled schedules)")
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    x = datagen.complex_data(100, (n,))
    y = fftm.ccfft(amd, x.copy(), fftm.View1(0, 1, n))
    _check("vs numpy fp64", y, np.fft.fft(x.astype(np.complex128)), 1e-5)


@pytest.mark.parametrize("m,n,major", [(2048, 2048, VSIP_ROW), (300, 1000, VSIP_ROW), (1000, 300, VSIP_COL),
                                       (4, 3 * 5 * 7 * 11, VSIP_ROW)])
def test_compiled_schedule_fp64(amd, monkeypatch, m, n, major):
    """FP64 plans compile the same schedule with double elements; 1e-12 to
    the FP64 oracle, 1e-13 to the generic kernel."""
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    x = datagen.complex_data(101, (m, n), dtype=np.complex128)
    v = fftm.dense_view(m, n)
    p = amd.ccfftmip_create_d(m, n, 1.0, VSIP_FFT_FWD, major, 0, 1)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p, "d"))
    amd.fftm_destroy_d(p)
    print("  passes", [(q["n"], q["radix"]) for q in d["passes"]], "compiled", d.get("compiled_schedule"),
          flush=True)
    assert d.get("compiled_schedule") is True
    for direction, scale in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / (n if major == VSIP_ROW else m))):
        _case("ccfftmip_d %s %dx%d dir %+d (compiled schedule)" % ("ROW" if major == VSIP_ROW else "COL", m, n,
                                                                  direction))
        y = fftm.ccfftm(amd, x.ravel(), v, major=major, direction=direction, scale=scale)
        _check("vs oracle", y, po.ccfftm(x.ravel(), v, major=major, direction=direction, scale=scale), 1e-12)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.ccfftm(amd, x.ravel(), v, major=major)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("vs generic kernel", fftm.ccfftm(amd, x.ravel(), v, major=major), g, 1e-13)


@pytest.mark.parametrize("m,n,major", [(300, 2000, VSIP_ROW), (2000, 300, VSIP_COL), (64, 2 * 2187, VSIP_ROW),
                                       (1024, 8192, VSIP_COL)])
def test_compiled_schedule_real_to_complex(amd, monkeypatch, m, n, major):
    """rcfftmop plans compile the real-input schedule (pack on load, combine
    in LDS); against the oracle (1e-5) and the generic kernel (1e-6)."""
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    x = datagen.real_data(102, (m, n))
    if major == VSIP_ROW:
        xv, zv, zshape = fftm.View(0, n, m, 1, n), fftm.View(0, n // 2 + 1, m, 1, n // 2 + 1), (m, n // 2 + 1)
    else:
        xv, zv, zshape = fftm.View(0, n, m, 1, n), fftm.View(0, n, m // 2 + 1, 1, n), (m // 2 + 1, n)
    p = amd.rcfftmop_create_f(m, n, 1.0, major, 0, 1)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    print("  passes", [(q["n"], q["radix"]) for q in d["passes"]], "compiled", d.get("compiled_schedule"),
          flush=True)
    pow2_row = major == VSIP_ROW and (n // 2) & (n // 2 - 1) == 0
    assert d.get("compiled_schedule") is (not pow2_row)
    _case("rcfftmop_f %s %dx%d (compiled schedule)" % ("ROW" if major == VSIP_ROW else "COL", m, n))
    z0 = np.zeros(zshape[0] * zshape[1], np.complex64)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, major=major, scale=0.5)
    _check("vs oracle", z, po.rcfftm(x.ravel(), xv, z0, zv, major, 0.5), 1e-5)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, major=major, scale=0.5)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("vs generic kernel", z, g, 1e-6)


@pytest.mark.parametrize("m,n,ntimes,want", [(16, 1000, 0, True), (16, 1000, 1, False), (16, 1000, 1 << 20, True),
                                             (8, 243, 0, False), (64, 2187, 100, False)])
def test_compile_rule_follows_ntimes(amd, monkeypatch, m, n, ntimes, want):
    """Without VSIP_AMD_JIT_MIN_ELEMS a plan compiles its schedule when its
    expected uses repay it: ntimes 0 (VSIPL: unbounded uses) from 2^12
    elements, ntimes > 0 when ntimes x elements >= 2^30."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    _case("ccfftmip_f ROW %dx%d ntimes %d: compiled %s" % (m, n, ntimes, want))
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, VSIP_ROW, ntimes, 0)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    assert bool(d.get("compiled_schedule")) is want
    x = datagen.complex_data(96, (m, n))
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v), 1e-5)


_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import load, VSIP_ROW, VSIP_FFT_FWD
amd = load()
m, n = int(sys.argv[2]), int(sys.argv[3])
p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, VSIP_ROW, 0, 0)
assert p, amd.last_error()
d = json.loads(amd.describe(p))
amd.fftm_destroy_f(p)
y = fftm.ccfftm(amd, datagen.complex_data(97, (m, n)).ravel(), fftm.dense_view(m, n))
np.save(sys.argv[4], y)
print(json.dumps({"compiled": d.get("compiled_schedule")}), flush=True)
"""


def _child_plan(cache_dir, m, n, out):
    """One plan in a fresh process (an in-process cache would hide the disk)."""
    env = dict(os.environ, VSIP_AMD_JIT_CACHE_DIR=str(cache_dir))
    env.pop("VSIP_AMD_JIT_MIN_ELEMS", None)
    env.pop("VSIP_AMD_JIT", None)
    r = subprocess.run([sys.executable, "-c", _CHILD, _ROOT, str(m), str(n), str(out)], env=env,
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1]), np.load(out)


def test_code_object_cache_on_disk(tmp_path):
    """VSIP_AMD_JIT_CACHE_DIR: the first process compiles and writes the
    schedule's code object; the next loads it (the file is not rewritten);
    a damaged file is compiled anew and replaced.  Every run matches the
    oracle."""
    m, n = 3, 5005                       # [13, 11, 7, 5]: a schedule no other test compiles
    want = po.ccfftm(datagen.complex_data(97, (m, n)).ravel(), fftm.dense_view(m, n))
    cache = tmp_path / "cache"
    cache.mkdir()
    _case("code object cache: first process compiles")
    d, y = _child_plan(cache, m, n, tmp_path / "y1.npy")
    assert d["compiled"] is True
    _check("vs oracle", y, want, 1e-5)
    files = sorted(cache.glob("vsip_amd_*.co"))
    assert len(files) == 1 and files[0].read_bytes().startswith(b"vsip_amd jit 1 ")
    ino = files[0].stat().st_ino
    _case("code object cache: second process loads the file")
    d, y2 = _child_plan(cache, m, n, tmp_path / "y2.npy")
    assert d["compiled"] is True and files[0].stat().st_ino == ino
    np.testing.assert_array_equal(y2, y)
    _case("code object cache: a damaged file is compiled anew")
    files[0].write_bytes(b"vsip_amd jit 1 nothing\n" + b"\0" * 256)
    d, y3 = _child_plan(cache, m, n, tmp_path / "y3.npy")
    assert d["compiled"] is True
    np.testing.assert_array_equal(y3, y)
    data = files[0].read_bytes()
    assert data.startswith(b"vsip_amd jit 1 ") and b"\x7fELF" in data[:4096]
    assert not list(cache.glob("*.tmp.*"))


@pytest.mark.parametrize("m,n", [(64, 1024), (300, 256), (100, 512), (16, 2048), (1000, 1024), (5, 2048)])
def test_small_power_of_two_rows_both_kernels(amd, monkeypatch, m, n):
    """Complex rows of 256..2048 in batches below the pipelined kernel's
    compile their schedule (it beats the one-group power-of-two kernel);
    with VSIP_AMD_JIT=0 they run on the one-group kernel.  Both against the
    oracle (1e-5) and each other (1e-6)."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    x = datagen.complex_data(98, (m, n))
    v = fftm.dense_view(m, n)
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, VSIP_ROW, 0, 0)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    assert d.get("compiled_schedule") is True
    _case("ccfftmip_f ROW %dx%d inverse (compiled schedule)" % (m, n))
    y = fftm.ccfftm(amd, x.ravel(), v, direction=VSIP_FFT_INV, scale=1.0 / n)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v, direction=VSIP_FFT_INV, scale=1.0 / n), 1e-5)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, VSIP_ROW, 0, 0)
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    assert d.get("compiled_schedule") is False
    _case("ccfftmip_f ROW %dx%d inverse (one-group power-of-two kernel)" % (m, n))
    g = fftm.ccfftm(amd, x.ravel(), v, direction=VSIP_FFT_INV, scale=1.0 / n)
    _check("vs oracle", g, po.ccfftm(x.ravel(), v, direction=VSIP_FFT_INV, scale=1.0 / n), 1e-5)
    _check("compiled vs power-of-two kernel", y, g, 1e-6)


def test_large_power_of_two_rows_stay_on_their_kernels(amd, monkeypatch):
    """4096-point rows and batches that fill the pipelined kernel never
    compile a schedule."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    for m, n in ((64, 4096), (8192, 1024), (4096, 2048), (32768, 256)):
        p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, VSIP_ROW, 0, 0)
        assert p, amd.last_error()
        d = json.loads(amd.describe(p))
        amd.fftm_destroy_f(p)
        assert d.get("compiled_schedule") is False, (m, n)


@pytest.mark.parametrize("m,n", [(64, 1024), (16, 4096), (100, 512), (1000, 2048)])
def test_small_power_of_two_real_rows_both_kernels(amd, monkeypatch, m, n):
    """rcfftmop rows of 512..4096 reals in batches below the pipelined R2C
    kernel's compile their schedule; with VSIP_AMD_JIT=0 they run on the
    one-group real kernel.  Both against the oracle (1e-5) and each other
    (1e-6)."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    x = datagen.real_data(99, (m, n))
    h = n // 2 + 1
    xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, h)
    z0 = np.zeros(m * h, np.complex64)
    p = amd.rcfftmop_create_f(m, n, 1.0, VSIP_ROW, 0, 0)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    assert d.get("compiled_schedule") is True
    _case("rcfftmop_f ROW %dx%d (compiled schedule)" % (m, n))
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, scale=0.5)
    _check("vs oracle", z, po.rcfftm(x.ravel(), xv, z0, zv, 0, 0.5), 1e-5)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    _case("rcfftmop_f ROW %dx%d (one-group real kernel)" % (m, n))
    g = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, scale=0.5)
    _check("vs oracle", g, po.rcfftm(x.ravel(), xv, z0, zv, 0, 0.5), 1e-5)
    _check("compiled vs real kernel", z, g, 1e-6)


@pytest.mark.parametrize("m,n,major,prec,rstride", [(300, 2000, VSIP_ROW, 4, 0), (2000, 300, VSIP_COL, 4, 0),
                                                    (64, 2 * 2187, VSIP_ROW, 4, 0), (64, 1024, VSIP_ROW, 4, 0),
                                                    (512, 4374, VSIP_ROW, 4, 0), (1000, 64, VSIP_COL, 4, 0),
                                                    (30, 1000, VSIP_ROW, 4, 1), (40, 2000, VSIP_ROW, 8, 0),
                                                    (600, 48, VSIP_COL, 8, 0)])
def test_compiled_schedule_complex_to_real(amd, monkeypatch, m, n, major, prec, rstride):
    """crfftmop plans compile the complex-to-real schedule (the n + 1 inputs
    combined into LDS, real samples stored in pairs or one by one); on
    non-Hermitian input against the oracle (1e-5 / 1e-12) and the generic
    kernel (1e-6 / 1e-13).  rstride 1 pads the real rows to an odd stride,
    which the compiled kernel does not take (the generic kernel runs it)."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    ctype, rtype = (np.complex64, np.float32) if prec == 4 else (np.complex128, np.float64)
    tol, tol_g = (1e-5, 1e-6) if prec == 4 else (1e-12, 1e-13)
    L = n if major == VSIP_ROW else m
    if major == VSIP_ROW:
        zv = fftm.View(0, n // 2 + 1, m, 1, n // 2 + 1)
        zlen = m * (n // 2 + 1)
    else:
        zv = fftm.View(0, n, m // 2 + 1, 1, n)
        zlen = (m // 2 + 1) * n
    xv = fftm.View(0, n + rstride, m, 1, n)
    z = datagen.complex_data(103, (zlen,), dtype=ctype)
    x0 = np.zeros(fftm.block_len(xv), rtype)
    create = amd.crfftmop_create_f if prec == 4 else amd.crfftmop_create_d
    p = create(m, n, 1.0 / L, major, 0, 0)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p, "f" if prec == 4 else "d"))
    (amd.fftm_destroy_f if prec == 4 else amd.fftm_destroy_d)(p)
    assert d.get("compiled_schedule") is True
    _case("crfftmop_%s %s %dx%d%s (compiled schedule)" % ("f" if prec == 4 else "d",
                                                          "ROW" if major == VSIP_ROW else "COL", m, n,
                                                          ", odd real row stride" if rstride else ""))
    x = fftm.crfftm(amd, z, zv, x0, xv, major=major, scale=1.0 / L)
    _check("vs oracle", x, po.crfftm(z, zv, x0, xv, major, 1.0 / L), tol)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.crfftm(amd, z, zv, x0, xv, major=major, scale=1.0 / L)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("vs generic kernel", x, g, tol_g)


@pytest.mark.parametrize("major,m,n", [(VSIP_ROW, 300, 1000), (VSIP_COL, 300, 1000), (VSIP_COL, 1920, 64)])
def test_other_layout_compiles_at_first_launch(amd, monkeypatch, major, m, n):
    """Column-major storage gives a COL plan contiguous columns and a ROW
    plan rows across the storage: layouts the plan does not compile at
    creation.  The first launch compiles them; every call matches the oracle
    (1e-5), the calls match each other bit for bit and the generic kernel
    to 1e-6."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    v = fftm.View(0, 1, m, m, n)
    x = datagen.complex_data(104, (m * n,))
    _case("ccfftmip_f %s %dx%d on column-major storage" % ("ROW" if major == VSIP_ROW else "COL", m, n))
    want = po.ccfftm(x, v, major=major)
    y1 = fftm.ccfftm(amd, x, v, major=major)
    _check("first call vs oracle", y1, want, 1e-5)
    y2 = fftm.ccfftm(amd, x, v, major=major)
    np.testing.assert_array_equal(y2, y1)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.ccfftm(amd, x, v, major=major)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("vs generic kernel", y1, g, 1e-6)


@pytest.mark.parametrize("major", [VSIP_ROW, VSIP_COL])
def test_one_plan_on_both_layouts(amd, monkeypatch, major):
    """One plan run in turn on a row-major view (the layout it compiled at
    creation) and a column-major view (compiled at its first launch): every
    call matches the oracle (1e-5) and the calls on each view repeat bit for
    bit."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    m, n = 240, 1000
    xr = datagen.complex_data(105, (m * n,))
    xc = datagen.complex_data(106, (m * n,))
    vr, vc = fftm.View(0, n, m, 1, n), fftm.View(0, 1, m, m, n)
    want_r = po.ccfftm(xr, vr, major=major)
    want_c = po.ccfftm(xc, vc, major=major)
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, major, 0, 0)
    assert p, amd.last_error()
    br = amd.cblockcreate_f(m * n, 0)
    bc = amd.cblockcreate_f(m * n, 0)
    hr = amd.cmbind_f(br, *vr)
    hc = amd.cmbind_f(bc, *vc)
    first = {}
    for it in range(3):
        for name, h, x, v, want in (("row-major", hr, xr, vr, want_r), ("column-major", hc, xc, vc, want_c)):
            _case("one %s plan, call %d on the %s view" % ("ROW" if major == VSIP_ROW else "COL", it, name))
            xx = np.ascontiguousarray(x)
            blk = br if h is hr else bc
            whole = amd.cmbind_f(blk, 0, 1, 1, 1, m * n)
            amd.cmcopyfrom_user_f(xx.ctypes.data, None, VSIP_ROW, whole)
            amd.ccfftmip_f(p, h)
            got = np.empty(m * n, np.complex64)
            amd.cmcopyto_user_f(whole, VSIP_ROW, got.ctypes.data, None)
            amd.cmdestroy_f(whole)
            _check("vs oracle", got, want, 1e-5)
            if name in first:
                np.testing.assert_array_equal(got, first[name])
            first[name] = got
    assert not amd.last_error(), amd.last_error()
    amd.fftm_destroy_f(p)
    amd.cmdestroy_f(hr)
    amd.cmdestroy_f(hc)
    amd.cblockdestroy_f(br)
    amd.cblockdestroy_f(bc)


@pytest.mark.parametrize("m,n,major,op,prec", [(64, 1000, VSIP_ROW, False, 4), (300, 2187, VSIP_ROW, True, 4),
                                               (1920, 100, VSIP_COL, False, 4), (256, 1024, VSIP_ROW, False, 4),
                                               (1000, 64, VSIP_COL, True, 4), (40, 2000, VSIP_ROW, False, 8),
                                               (500, 48, VSIP_COL, True, 8)])
def test_split_storage_compiled(amd, monkeypatch, m, n, major, op, prec):
    """Split complex blocks (real and imaginary arrays) run the compiled
    schedule (compiled at the first launch for the split layout): against the
    oracle (1e-5 / 1e-12), the same transform on interleaved blocks (1e-6 /
    1e-13) and the generic kernel on the split blocks."""
    monkeypatch.delenv("VSIP_AMD_JIT_MIN_ELEMS", raising=False)
    monkeypatch.delenv("VSIP_AMD_JIT", raising=False)
    ctype = np.complex64 if prec == 4 else np.complex128
    tol, tol2 = (1e-5, 1e-6) if prec == 4 else (1e-12, 1e-13)
    x = datagen.complex_data(107, (m * n,), dtype=ctype)
    v = fftm.dense_view(m, n)
    _case("ccfftm%s_%s %s %dx%d on split blocks" % ("op" if op else "ip", "f" if prec == 4 else "d",
                                                    "ROW" if major == VSIP_ROW else "COL", m, n))
    kw = dict(major=major, direction=VSIP_FFT_INV, scale=0.5)
    if op:
        y0 = np.zeros(m * n, ctype)
        want = po.ccfftm(x, v, out_blk=y0, out_view=v, **kw)
        xr, y = fftm.ccfftm(amd, x, v, out_blk=y0, out_view=v, split=True, **kw)
        np.testing.assert_array_equal(xr, x)
        _, yi = fftm.ccfftm(amd, x, v, out_blk=y0, out_view=v, **kw)
    else:
        want = po.ccfftm(x, v, **kw)
        y = fftm.ccfftm(amd, x, v, split=True, **kw)
        yi = fftm.ccfftm(amd, x, v, **kw)
    _check("split vs oracle", y, want, tol)
    _check("split vs interleaved", y, yi, tol2)
    monkeypatch.setenv("VSIP_AMD_JIT", "0")
    g = fftm.ccfftm(amd, x, v, out_blk=np.zeros(m * n, ctype), out_view=v, split=True, **kw)[1] if op else \
        fftm.ccfftm(amd, x, v, split=True, **kw)
    monkeypatch.delenv("VSIP_AMD_JIT")
    _check("split compiled vs generic", y, g, tol2)
