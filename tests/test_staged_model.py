"""CPU checks of the staged-transform arithmetic (no device needed).

The staged plans (jvsip_amd/csrc/host/va_stage.c, kernels in
jvsip_amd/csrc/hip/fftm_staged.hip) compute an n-point DFT of any length by
Bluestein's chirp-z on scratch rows of a smooth length sm, and real
transforms by packing/combining around a complex FFT.  Here:
  - the library's host FFT (va_host_fft, which builds the chirp filter's
    spectrum in double precision) is checked against numpy.fft, an
    independent routine, at smooth, prime and mixed lengths;
  - the convolution length rule is checked (>= 2n - 1, 2^a 3^b 5^c);
  - a numpy model of the PRE / FFT / MUL / FFT / POST passes, with the same
    index, conjugation and chirp rules as the kernels and the filter table
    from va_host_fft, reproduces numpy.fft's DFT, inverse DFT, real forward
    and (non-Hermitian) real inverse as the oracle defines them -- so a
    wrong sign, a missing index reversal or a misplaced conjugation in that
    scheme fails here before any kernel runs.
"""
import ctypes

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from oracle import pyoracle as po


@pytest.fixture(scope="module")
def so():
    from jvsip_amd.vsip import LIB_PATH
    d = ctypes.CDLL(LIB_PATH)
    d.va_host_fft.argtypes = [ctypes.c_long, ctypes.c_void_p]
    d.va_host_fft.restype = None
    d.va_bluestein_length.argtypes = [ctypes.c_long]
    d.va_bluestein_length.restype = ctypes.c_long
    return d


def _host_fft(so, x):
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    so.va_host_fft(len(y), y.ctypes.data)
    return y


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 17, 97, 210, 256, 540, 1000, 1009, 4096, 8640, 36000])
def test_host_fft_vs_numpy(so, n):
    x = datagen.complex_data(30, (n,), dtype=np.complex128)
    err = po.rel_l2(_host_fft(so, x), np.fft.fft(x))
    assert err < 1e-13, (n, err)


def test_bluestein_length_rule(so):
    def smooth(m):
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        return m == 1
    for n in list(range(1, 300)) + [1009, 4099, 17408, 65537, 1 << 20]:
        m = so.va_bluestein_length(n)
        assert m >= 2 * n - 1 and smooth(m), (n, m)
        best = min(k for k in range(max(1, 2 * n - 1), 4 * n + 2) if smooth(k))
        p2 = 1 << (2 * n - 2).bit_length() if n > 1 else 1
        assert m == (p2 if 4 * p2 <= 5 * best else best), (n, m, best, p2)


def _chirp(n):
    j = np.arange(n, dtype=np.int64)
    return np.exp(-1j * np.pi * ((j * j) % (2 * n)) / n)


def _model(so, kind, data, n, conj=False, scale=1.0):
    """numpy restatement of the kernel passes for one vector (fftm_staged.hip)."""
    sm = so.va_bluestein_length(n)
    a = _chirp(n)
    b = np.zeros(sm, np.complex128)
    b[:n] = np.conj(a)
    b[sm - np.arange(1, n)] = np.conj(a[1:])
    table = _host_fft(so, b) / sm
    w = np.zeros(sm, np.complex128)
    if kind == "r2c":
        y = data[0::2] + 1j * data[1::2]
    elif kind == "c2r":
        Z = data
        k = np.arange(n)
        F, Bc = Z[k], np.conj(Z[n - k])
        e = np.exp(1j * np.pi * k / n)
        y = np.conj((F + Bc) + 1j * e * (F - Bc))
    else:
        y = np.conj(data) if conj else data.astype(np.complex128)
    w[:n] = y * a                                            # PRE
    w = np.fft.fft(w)                                        # FFT (forward)
    w = w * table                                            # MUL
    w = np.fft.fft(w)                                        # FFT (forward)
    k = np.arange(n)
    Y = a * w[(sm - k) % sm]                                 # POST: reverse + chirp
    if kind == "r2c":
        kk = np.arange(n + 1)
        F = Y[kk % n]
        Bc = np.conj(Y[(n - kk) % n])
        e = np.exp(-1j * np.pi * kk / n)
        return 0.5 * scale * ((F + Bc) - 1j * e * (F - Bc))
    if kind == "c2r":
        out = np.empty(2 * n)
        out[0::2] = scale * Y.real
        out[1::2] = -scale * Y.imag
        return out
    return scale * (np.conj(Y) if conj else Y)


@pytest.mark.parametrize("n", [1, 2, 17, 38, 97, 263, 1009])
def test_bluestein_model_c2c(so, n):
    x = datagen.complex_data(31, (n,), dtype=np.complex128)
    assert po.rel_l2(_model(so, "c2c", x, n), np.fft.fft(x)) < 1e-12
    inv = _model(so, "c2c", x, n, conj=True, scale=0.5)
    assert po.rel_l2(inv, 0.5 * n * np.fft.ifft(x)) < 1e-12
    # the same vector through the oracle's definition (sign +1, scale)
    want = po.ccfftm(x, fftm.dense_view(1, n), direction=1, scale=0.5)
    assert po.rel_l2(inv, want) < 1e-12


@pytest.mark.parametrize("n", [17, 101, 1009])
def test_bluestein_model_real(so, n):
    x = datagen.real_data(32, (2 * n,)).astype(np.float64)
    z = _model(so, "r2c", x, n, scale=2.0)
    assert po.rel_l2(z, 2.0 * np.fft.rfft(x)) < 1e-12
    zz = datagen.complex_data(33, (n + 1,), dtype=np.complex128)   # not Hermitian
    got = _model(so, "c2r", zz, n, scale=0.25)
    want = po.crfftm(zz.astype(np.complex128), fftm.dense_view(1, n + 1), np.zeros(2 * n),
                     fftm.dense_view(1, 2 * n), scale=0.25)
    assert po.rel_l2(got, want) < 1e-12


def test_model_detects_planted_mistakes(so):
    """Each plausible slip in the scheme breaks the model's agreement."""
    n = 97
    x = datagen.complex_data(34, (n,), dtype=np.complex128)
    want = np.fft.fft(x)
    sm = so.va_bluestein_length(n)
    a = _chirp(n)
    b = np.zeros(sm, np.complex128)
    b[:n] = np.conj(a)
    b[sm - np.arange(1, n)] = np.conj(a[1:])
    table = _host_fft(so, b) / sm
    k = np.arange(n)

    def run(a_pre, tab, rev):
        w = np.zeros(sm, np.complex128)
        w[:n] = x * a_pre
        w = np.fft.fft(np.fft.fft(w) * tab)
        return a * (w[(sm - k) % sm] if rev else w[k])

    assert po.rel_l2(run(a, table, True), want) < 1e-12
    assert po.rel_l2(run(np.conj(a), table, True), want) > 1e-2          # chirp sign
    assert po.rel_l2(run(a, table, False), want) > 1e-2                   # no index reversal
    assert po.rel_l2(run(a, np.conj(table), True), want) > 1e-2           # filter conjugated
