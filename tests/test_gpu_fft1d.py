"""GPU parity of the 1-D FFT (vsip_ccfftip/op, vsip_rcfftop, vsip_crfftop,
_f and _d) through the C ABI, against the oracle (the multiple FFT of the
1 x n view, which the reference's own 1-D path equals: tests/test_oracle.py),
jvsip's 1-D calls, and the closed form of the reference's own 1-D test
(c_VSIP_testing/rcfftop_f.h:40-60).

Tolerances: FP32 1e-5, FP64 1e-12 relative L2.
"""
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_FFT_FWD, VSIP_FFT_INV, fft_attr_f
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


@pytest.mark.parametrize("n", [8, 22, 1000, 1001, 4096, 4099, 65536, 1 << 20])
def test_ccfft_vs_oracle_and_jvsip(amd, ref, n):
    x = datagen.complex_data(80, (n,))
    v = fftm.View1(0, 1, n)
    for d, s in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / n)):
        _case("ccfftip_f n=%d dir %+d" % (n, d))
        y = fftm.ccfft(amd, x.copy(), v, direction=d, scale=s)
        if n <= 65536 and n != 4099:
            _check("vs oracle", y, po.ccfftm(x, fftm.as_matrix(v), 0, d, s), TOL_F)
        else:
            want = (np.fft.fft(x.astype(np.complex128)) if d < 0 else np.fft.ifft(x.astype(np.complex128)) * n) * s
            _check("vs numpy fp64", y, want, TOL_F)
        if n <= 4096:
            _check("vs jvsip", y, fftm.ccfft(ref, x.copy(), v, direction=d, scale=s), TOL_F)


def test_ccfftop_strided_and_reversed(amd):
    n = 2048
    _case("ccfftop_f n=%d, input stride 2, output stride -1" % n)
    x = datagen.complex_data(81, (2 * n,))
    xv = fftm.View1(0, 2, n)
    yv = fftm.View1(n - 1, -1, n)
    xr, y = fftm.ccfft(amd, x, xv, out_blk=np.zeros(n, np.complex64), out_view=yv)
    np.testing.assert_array_equal(xr, x)
    want = po.ccfftm(x, fftm.as_matrix(xv), 0, -1, 1.0, out_blk=np.zeros(n, np.complex64),
                     out_view=fftm.as_matrix(yv))
    _check("vs oracle", y, want, TOL_F)


@pytest.mark.parametrize("n", [8, 16, 22, 24])
def test_rcfftop_ramp_reference_test(amd, n):
    """c_VSIP_testing/rcfftop_f.h:40-60: the ramp 0..n-1 with scale 1/n."""
    _case("rcfftop_f ramp n=%d scale 1/n" % n)
    h = n // 2 + 1
    z = fftm.rcfft(amd, np.arange(n, dtype=np.float32), fftm.View1(0, 1, n), np.zeros(h, np.complex64),
                   fftm.View1(0, 1, h), scale=1.0 / n)
    k = np.arange(1, n // 2)
    want = np.concatenate([[(n - 1) / 2.0], -0.5 + 0.5j / np.tan(np.pi * k / n), [-0.5]])
    assert np.max(np.abs(z - want)) < 1e-3          # the reference test's own criterion
    _check("vs closed form", z, want, TOL_F)


@pytest.mark.parametrize("n", [24, 2048, 2 * 1009, 65536])
def test_real_fft1d(amd, ref, n):
    h = n // 2 + 1
    _case("rcfftop_f / crfftop_f n=%d" % n)
    r = datagen.real_data(82, (n,))
    z = fftm.rcfft(amd, r, fftm.View1(0, 1, n), np.zeros(h, np.complex64), fftm.View1(0, 1, h))
    _check("rc vs numpy fp64", z, np.fft.rfft(r.astype(np.float64)), TOL_F)
    zz = datagen.complex_data(83, (h,))                  # not Hermitian
    xb = fftm.crfft(amd, zz, fftm.View1(0, 1, h), np.zeros(n, np.float32), fftm.View1(0, 1, n), scale=0.5)
    _check("cr vs oracle", xb, po.crfftm(zz, fftm.dense_view(1, h), np.zeros(n), fftm.dense_view(1, n), scale=0.5),
           TOL_F)
    if n <= 4096:
        _check("cr vs jvsip", xb, fftm.crfft(ref, zz, fftm.View1(0, 1, h), np.zeros(n, np.float32),
                                            fftm.View1(0, 1, n), scale=0.5), TOL_F)
    back = fftm.crfft(amd, z, fftm.View1(0, 1, h), np.zeros(n, np.float32), fftm.View1(0, 1, n), scale=1.0 / n)
    _check("round trip", back, r, 1e-6)


@pytest.mark.parametrize("n", [64, 1000, 257, 16384])
def test_ccfft_d(amd, ref, n):
    _case("ccfftip_d n=%d" % n)
    x = datagen.complex_data(84, (n,), dtype=np.complex128)
    v = fftm.View1(0, 1, n)
    y = fftm.ccfft(amd, x.copy(), v)
    _check("vs numpy fp64", y, np.fft.fft(x), TOL_D)
    if n <= 1000:
        _check("vs jvsip _d", y, fftm.ccfft(ref, x.copy(), v), TOL_D)


def test_fft_getattr(amd):
    a = fft_attr_f()
    for create, args, want in (("ccfftip_create_f", (24, 0.5, VSIP_FFT_INV, 0, 1), (24, 24, 0, 1)),
                               ("ccfftop_create_f", (24, 0.5, VSIP_FFT_FWD, 0, 1), (24, 24, 1, -1)),
                               ("rcfftop_create_f", (24, 0.5, 0, 1), (24, 13, 1, -1)),
                               ("crfftop_create_f", (24, 0.5, 0, 1), (13, 24, 1, 1))):
        p = getattr(amd, create)(*args)
        assert p, amd.last_error()
        amd.fft_getattr_f(p, a)
        assert (a.input, a.output, a.place, a.dir) == want and abs(a.scale - 0.5) < 1e-7, create
        assert amd.fft_destroy_f(p) == 0
    assert not amd.rcfftop_create_f(23, 1.0, 0, 1)          # odd real length refused
    amd.amd_clear_error()
