"""GPU parity of staged transforms through the C ABI.

Staged plans cover the lengths the single-pass kernels and the two-pass split
do not: lengths with a prime factor above 13 (Bluestein's chirp-z on smooth
scratch rows) and real transforms whose half length needs two passes.  The
reference transforms every length (its remainder DFT,
VI_fftm_building_blocks_f.h:95-155), so these must agree with the oracle,
with jvsip where its O(L^2) remainder DFT is cheap enough to run, and, for
long prime lengths where the oracle's O(p^2) prime step is too slow, with
numpy.fft in double precision (pinned against the oracle on the CPU).

Tolerances as for every other plan: FP32 1e-5 relative L2, FP64 1e-12.
"""
import json
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL_F = 1e-5
TOL_D = 1e-12


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _check(name, got, want, tol):
    err = po.rel_l2(got, want)
    print("  %s: rel L2 %.3e" % (name, err), flush=True)
    assert err < tol, "%s: rel L2 %.3e >= %.1e" % (name, err, tol)
    return err


def _np_dft(x, direction, scale, axis=1):
    x = x.astype(np.complex128)
    y = np.fft.fft(x, axis=axis) if direction < 0 else np.fft.ifft(x, axis=axis) * x.shape[axis]
    return y * scale


def _describe(lib, m, n, prec=4, major=VSIP_ROW):
    t = "f" if prec == 4 else "d"
    p = getattr(lib, "ccfftmip_create_" + t)(m, n, 1.0, VSIP_FFT_FWD, major, 0, 1)
    assert p, lib.last_error()
    d = json.loads(lib.describe(p, t))
    getattr(lib, "fftm_destroy_" + t)(p)
    return d


@pytest.mark.parametrize("n", [17, 19 * 2, 97, 257, 1009, 17 * 64])
def test_bluestein_row_vs_oracle_and_jvsip(amd, ref, n):
    m = 5
    d = _describe(amd, m, n)
    print("  schedule", d, flush=True)
    assert d.get("staged") == "bluestein"
    x = datagen.complex_data(11, (m, n))
    v = fftm.dense_view(m, n)
    for direction, scale in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / n)):
        _case("ccfftmip_f ROW %dx%d dir %+d (Bluestein)" % (m, n, direction))
        y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=scale)
        _check("vs oracle", y, po.ccfftm(x.ravel(), v, direction=direction, scale=scale), TOL_F)
        _check("vs jvsip", y, fftm.ccfftm(ref, x.ravel(), v, direction=direction, scale=scale), TOL_F)


@pytest.mark.parametrize("n", [4099, 17 * 1024, 65537])
def test_bluestein_long(amd, n):
    m = 3
    d = _describe(amd, m, n)
    print("  schedule", d, flush=True)
    x = datagen.complex_data(12, (m, n))
    v = fftm.dense_view(m, n)
    for direction in (VSIP_FFT_FWD, VSIP_FFT_INV):
        _case("ccfftmip_f ROW %dx%d dir %+d (Bluestein, long)" % (m, n, direction))
        y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=0.5)
        _check("vs numpy fp64", y.reshape(m, n), _np_dft(x, direction, 0.5), TOL_F)
    if n < 5000:
        _check("vs oracle", fftm.ccfftm(amd, x.ravel(), v), po.ccfftm(x.ravel(), v), TOL_F)


def test_bluestein_col_strided_out_of_place(amd, ref):
    # 101-point columns of a view whose rows lie twice the row length apart
    m, n = 101, 24
    _case("ccfftmop_f COL %dx%d, parent rows 2x (Bluestein)" % (m, n))
    x = datagen.complex_data(13, (m * 2 * n,))
    xv = fftm.View(0, 2 * n, m, 1, n)
    y0 = np.zeros(m * 2 * n, np.complex64)
    yv = fftm.View(n, 2 * n, m, 1, n)
    xr, y = fftm.ccfftm(amd, x, xv, major=VSIP_COL, out_blk=y0, out_view=yv, direction=VSIP_FFT_INV,
                        scale=2.0)
    np.testing.assert_array_equal(xr, x)
    want = po.ccfftm(x, xv, major=VSIP_COL, out_blk=y0, out_view=yv, direction=VSIP_FFT_INV, scale=2.0)
    _check("vs oracle", y, want, TOL_F)
    _, yr = fftm.ccfftm(ref, x, xv, major=VSIP_COL, out_blk=y0, out_view=yv, direction=VSIP_FFT_INV,
                        scale=2.0)
    _check("vs jvsip", y, yr, TOL_F)


def test_bluestein_split_complex(amd):
    _case("ccfftmip_f split block 6x127 (Bluestein)")
    x = datagen.complex_data(14, (6, 127))
    v = fftm.dense_view(6, 127)
    y = fftm.ccfftm(amd, x.ravel(), v, split=True)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v), TOL_F)


@pytest.mark.parametrize("m,n,major", [(4, 257, VSIP_ROW), (1031, 3, VSIP_COL), (2, 4099 * 2, VSIP_ROW)])
def test_bluestein_fp64(amd, ref, m, n, major):
    _case("ccfftmip_d %dx%d major %d (Bluestein)" % (m, n, major))
    x = datagen.complex_data(15, (m, n), dtype=np.complex128)
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v, major=major)
    axis = 1 if major == VSIP_ROW else 0
    _check("vs numpy fp64", y.reshape(m, n), _np_dft(x, VSIP_FFT_FWD, 1.0, axis), TOL_D)
    if n * m < 5000:
        _check("vs oracle", y, po.ccfftm(x.ravel(), v, major=major), TOL_D)
        _check("vs jvsip _d", y, fftm.ccfftm(ref, x.ravel(), v, major=major), TOL_D)


@pytest.mark.parametrize("m,n", [(6, 34), (5, 202), (3, 2 * 1009)])
def test_real_bluestein(amd, ref, m, n):
    _case("rcfftmop_f / crfftmop_f ROW %dx%d (Bluestein half length %d)" % (m, n, n // 2))
    x = datagen.real_data(16, (m, n))
    xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, n // 2 + 1)
    z0 = np.zeros(m * (n // 2 + 1), np.complex64)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv)
    _check("rc vs oracle", z, po.rcfftm(x.ravel(), xv, z0, zv), TOL_F)
    _check("rc vs jvsip", z, fftm.rcfftm(ref, x.ravel(), xv, z0, zv), TOL_F)
    zz = datagen.complex_data(17, (m * (n // 2 + 1),))      # not Hermitian: the literal combine
    x0 = np.zeros(m * n, np.float32)
    xb = fftm.crfftm(amd, zz, zv, x0, xv, scale=0.5)
    _check("cr vs oracle", xb, po.crfftm(zz, zv, x0, xv, scale=0.5), TOL_F)
    _check("cr vs jvsip", xb, fftm.crfftm(ref, zz, zv, x0, xv, scale=0.5), TOL_F)
    back = fftm.crfftm(amd, z, zv, x0, xv, scale=1.0 / n)
    _check("round trip", back, x.ravel(), 1e-6)


@pytest.mark.parametrize("m,n", [(3, 65536), (2, 40000)])
def test_real_two_pass(amd, m, n):
    _case("rcfftmop_f / crfftmop_f ROW %dx%d (staged around the two-pass split)" % (m, n))
    x = datagen.real_data(18, (m, n))
    xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, n // 2 + 1)
    z0 = np.zeros(m * (n // 2 + 1), np.complex64)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv)
    _check("rc vs numpy fp64", z.reshape(m, -1), np.fft.rfft(x.astype(np.float64), axis=1), TOL_F)
    zz = datagen.complex_data(19, (m * (n // 2 + 1),))
    x0 = np.zeros(m * n, np.float32)
    xb = fftm.crfftm(amd, zz, zv, x0, xv, scale=0.25)
    _check("cr vs oracle", xb, po.crfftm(zz, zv, x0, xv, scale=0.25), TOL_F)
    back = fftm.crfftm(amd, z, zv, x0, xv, scale=1.0 / n)
    _check("round trip", back, x.ravel(), 1e-6)


def test_real_col_bluestein(amd):
    m, n = 2 * 53, 20
    _case("rcfftmop_f / crfftmop_f COL %dx%d (Bluestein)" % (m, n))
    x = datagen.real_data(20, (m, n))
    xv = fftm.dense_view(m, n, VSIP_COL)
    zv = fftm.dense_view(m // 2 + 1, n, VSIP_COL)
    z0 = np.zeros((m // 2 + 1) * n, np.complex64)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, major=VSIP_COL)
    _check("rc vs oracle", z, po.rcfftm(x.ravel(), xv, z0, zv, major=VSIP_COL), TOL_F)
    xb = fftm.crfftm(amd, z, zv, np.zeros(m * n, np.float32), xv, major=VSIP_COL, scale=1.0 / m)
    _check("round trip", xb, x.ravel(), 1e-6)


def test_staged_slabs(amd, monkeypatch):
    """Scratch for 3 vectors at a time: 10 vectors run as 4 slabs, the last
    one partial, for a Bluestein and a long real plan."""
    n = 263
    sm = 540                                          # 2^2 3^3 5 >= 2 * 263 - 1
    monkeypatch.setenv("VSIP_AMD_STAGE_BYTES", str(3 * sm * 8))
    m = 10
    d = _describe(amd, m, n)
    print("  schedule", d, flush=True)
    assert d["chunk"] == 3 and d["scratch_length"] == sm
    _case("ccfftmip_f ROW %dx%d in slabs of 3" % (m, n))
    x = datagen.complex_data(21, (m, n))
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v, direction=VSIP_FFT_INV)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v, direction=VSIP_FFT_INV), TOL_F)
    _case("ccfftmip_f COL %dx%d in slabs of 3" % (n, m))
    y = fftm.ccfftm(amd, x.ravel(), fftm.dense_view(n, m), major=VSIP_COL)
    _check("vs oracle", y, po.ccfftm(x.ravel(), fftm.dense_view(n, m), major=VSIP_COL), TOL_F)
    nr = 65536
    monkeypatch.setenv("VSIP_AMD_STAGE_BYTES", str(2 * (nr // 2) * 8))
    m = 5
    _case("rcfftmop_f ROW %dx%d in slabs of 2" % (m, nr))
    x = datagen.real_data(22, (m, nr))
    xv, zv = fftm.dense_view(m, nr), fftm.dense_view(m, nr // 2 + 1)
    z = fftm.rcfftm(amd, x.ravel(), xv, np.zeros(m * (nr // 2 + 1), np.complex64), zv)
    _check("rc vs numpy fp64", z.reshape(m, -1), np.fft.rfft(x.astype(np.float64), axis=1), TOL_F)
