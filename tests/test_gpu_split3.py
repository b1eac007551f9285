"""GPU parity of the three-pass split (MN = n1 n2 n3, lengths from 2^21 by
default; VSIP_AMD_SPLIT3_MIN lowers the threshold so that shorter lengths
beyond one workgroup, from 16385, take the same path) against numpy's FP64 FFT (rel L2 <= 1e-5), for rows, columns,
strided out-of-place views, both directions, with and without the
compiled-schedule kernels, and for single long 1-D vectors.
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


def _npass(amd, m, n, major):
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, major, 0, 1)
    assert p, amd.last_error()
    d = json.loads(amd.describe(p))
    amd.fftm_destroy_f(p)
    print("  passes", [q["n"] for q in d["passes"]], "compiled", d.get("compiled_schedule"), flush=True)
    return d


def _want(x, axis, direction, scale):
    x = x.astype(np.complex128)
    n = x.shape[axis]
    return (np.fft.fft(x, axis=axis) if direction < 0 else np.fft.ifft(x, axis=axis) * n) * scale


@pytest.mark.parametrize("jit", ["1", "0"])
@pytest.mark.parametrize("m,n", [(3, 20480), (5, 30000), (2, 65536), (4, 3 * 5 * 7 * 256)])
def test_three_pass_rows(amd, monkeypatch, m, n, jit):
    monkeypatch.setenv("VSIP_AMD_SPLIT3_MIN", "4096")
    monkeypatch.setenv("VSIP_AMD_JIT", jit)
    monkeypatch.setenv("VSIP_AMD_JIT_MIN_ELEMS", "0")
    d = _npass(amd, m, n, VSIP_ROW)
    assert d["npass"] == 3 and (jit == "0" or d.get("compiled_schedule") is True)
    x = datagen.complex_data(130, (m, n))
    v = fftm.dense_view(m, n)
    for direction, scale in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / n)):
        _case("ccfftmip_f ROW %dx%d dir %+d three-pass (jit %s)" % (m, n, direction, jit))
        y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=scale).reshape(m, n)
        err = po.rel_l2(y, _want(x, 1, direction, scale))
        print("  rel L2 %.3e" % err, flush=True)
        assert err < 1e-5


def test_three_pass_columns_and_strided_out_of_place(amd, monkeypatch):
    monkeypatch.setenv("VSIP_AMD_SPLIT3_MIN", "4096")
    m, n = 32768, 3
    _case("ccfftmip_f COL %dx%d three-pass" % (m, n))
    assert _npass(amd, m, n, VSIP_COL)["npass"] == 3
    x = datagen.complex_data(131, (m, n))
    y = fftm.ccfftm(amd, x.ravel(), fftm.dense_view(m, n), major=VSIP_COL).reshape(m, n)
    assert po.rel_l2(y, _want(x, 0, -1, 1.0)) < 1e-5
    m, n = 3, 24576
    _case("ccfftmop_f ROW %dx%d, rows 2n apart in and out, inverse, three-pass" % (m, n))
    xb = datagen.complex_data(132, (m * 2 * n,))
    xv = fftm.View(0, 2 * n, m, 1, n)
    yv = fftm.View(n, 2 * n, m, 1, n)
    xr, yb = fftm.ccfftm(amd, xb, xv, out_blk=np.zeros(m * 2 * n, np.complex64), out_view=yv,
                         direction=VSIP_FFT_INV, scale=0.5)
    np.testing.assert_array_equal(xr, xb)
    got = yb.reshape(m, 2 * n)[:, n:]
    assert po.rel_l2(got, _want(xb.reshape(m, 2 * n)[:, :n], 1, 1, 0.5)) < 1e-5
    assert np.all(yb.reshape(m, 2 * n)[:, :n] == 0)


@pytest.mark.parametrize("n", [1 << 21, 1 << 22, 3 << 20, 1 << 24])
def test_three_pass_long_1d(amd, n):
    _case("ccfftip_f n=%d (three-pass by default)" % n)
    x = datagen.complex_data(133, (n,), dist="uniform")
    p = amd.ccfftip_create_f(n, 1.0, VSIP_FFT_FWD, 0, 1)
    assert p
    amd.fft_destroy_f(p)
    y = fftm.ccfft(amd, x.copy(), fftm.View1(0, 1, n))
    err = po.rel_l2(y, np.fft.fft(x.astype(np.complex128)))
    print("  rel L2 %.3e" % err, flush=True)
    assert err < 1e-5
