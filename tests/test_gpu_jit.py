"""GPU parity of the kernels compiled at plan creation (schedule-specialised
Stockham, fftm_static.hpp), through the C ABI.

With the size threshold at zero every eligible plan (FP32 complex rows, one
pass, not a power of two of the power-of-two kernel) compiles its schedule;
the plan must report the compiled kernel and its results must match the
oracle (rel L2 <= 1e-5) and the schedule-generic kernel (<= 1e-6; the two run
the same arithmetic).
"""
import json
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
