"""The correlation oracle (oracle/pyoracle.py: correlate1d, corr_divisors),
pinned independently of its own formulas:

  - biased output against the brute-force sum the reference's FFT steps
    reduce to (c_VSIP_src/vsip_ccorrelate1d_f.c:123-196 with N above
    m + n - 1, so the circular product never wraps):
        FULL y[i] = sum_j h[j] conj(x[i + j - (m - 1)]), x = 0 outside [0, n)
    SAME and MIN are the lags (m - 1)/2 .. and m - 1 .. of it; real
    correlation has no conjugate;
  - unbiased output: the reference's divisors are the number of samples of
    x that overlap h at each lag, min(i + 1, m, m + n - 1 - i) on the FULL
    lags (m <= n);
  - against the reference's own build (oracle/_ref), FP32 and FP64, real and
    complex, strided views, when it is built;
  - a planted mistake (the conjugate on h instead of x, a SAME offset off
    by one) fails the brute-force check.
"""
import numpy as np
import pytest

from jvsip_amd import fftm
from jvsip_amd.fftm import View1
from oracle import pyoracle as po

SHAPES = [(1, 1), (1, 7), (2, 2), (3, 3), (5, 12), (8, 8), (7, 20), (6, 21), (16, 100), (31, 64), (64, 65)]


def _brute_full(h, x):
    m, n = h.size, x.size
    xc = np.conj(x) if np.iscomplexobj(x) else x
    y = np.zeros(m + n - 1, np.result_type(h, x, np.float64))
    for i in range(m + n - 1):
        for j in range(m):
            k = i + j - (m - 1)
            if 0 <= k < n:
                y[i] += h[j] * xc[k]
    return y


def _window(full, m, n, support):
    if support == 0:
        return full
    if support == 1:
        return full[(m - 1) // 2:(m - 1) // 2 + n]
    return full[m - 1:n]


def _draw(rng, m, n, cplx):
    if cplx:
        return (rng.standard_normal(m) + 1j * rng.standard_normal(m),
                rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return rng.standard_normal(m), rng.standard_normal(n)


@pytest.mark.parametrize("cplx", [True, False])
@pytest.mark.parametrize("support", [0, 1, 2])
def test_biased_matches_brute_force(cplx, support):
    rng = np.random.default_rng(11 + support)
    for m, n in SHAPES:
        h, x = _draw(rng, m, n, cplx)
        got = po.correlate1d(h, x, support, 0)
        want = _window(_brute_full(h, x), m, n, support)
        assert got.shape == want.shape, (m, n, support)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(want).max()))
        assert np.iscomplexobj(got) == cplx


def _overlap(m, n, support):
    i = np.arange(m + n - 1)
    full = np.minimum(np.minimum(i + 1, m), m + n - 1 - i).astype(np.float64)
    return _window(full, m, n, support)


@pytest.mark.parametrize("support", [0, 1, 2])
def test_unbiased_divides_by_the_overlap(support):
    rng = np.random.default_rng(5)
    for m in range(1, 14):
        for n in range(m, m + 9):
            np.testing.assert_array_equal(po.corr_divisors(m, n, support), _overlap(m, n, support))
    for m, n in SHAPES:
        h, x = _draw(rng, m, n, True)
        np.testing.assert_allclose(po.correlate1d(h, x, support, 1) * _overlap(m, n, support),
                                   po.correlate1d(h, x, support, 0), rtol=1e-12, atol=1e-12)


def test_planted_mistakes_fail():
    rng = np.random.default_rng(3)
    h, x = _draw(rng, 7, 20, True)
    want = _brute_full(h, x)
    wrong_conj = np.conj(_brute_full(h, x))                          # conj on h, not x
    assert np.abs(po.correlate1d(h, x, 0, 0) - want).max() < 1e-10
    assert np.abs(wrong_conj - want).max() > 1e-3
    shifted = want[(7 - 1) // 2 + 1:(7 - 1) // 2 + 1 + 20]
    assert shifted.size < 20 or np.abs(shifted - po.correlate1d(h, x, 1, 0)).max() > 1e-3
    assert np.abs(_window(want, 7, 20, 1) - po.correlate1d(h, x, 1, 0)).max() < 1e-10


def test_fft_length():
    for m, n in [(1, 1), (3, 6), (4, 5), (5, 12), (1, 1024), (513, 1536)]:
        N = po.corr_length(m, n)
        assert N > m + n - 1 and N // 2 <= m + n - 1 and N & (N - 1) == 0


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("cplx", [True, False])
def test_oracle_matches_reference_build(ref, prec, cplx):
    """rrjudd/jvsip's own vsip_[c]correlate1d_[fd] (compiled from its sources)
    on strided views against the oracle: 1e-5 (FP32), 1e-12 (FP64)."""
    rng = np.random.default_rng(40 + prec + cplx)
    dt = (np.complex64 if prec == 4 else np.complex128) if cplx else (np.float32 if prec == 4 else np.float64)
    tol = 1e-5 if prec == 4 else 1e-12
    for m, n in SHAPES + [(100, 1000), (257, 300)]:
        for support in (0, 1, 2):
            for bias in (0, 1):
                h, x = _draw(rng, m, n, cplx)
                lag = m + n - 1 if support == 0 else n if support == 1 else n - m + 1
                hs, xs, ys = 1 + (m % 2), 2, 1 + (n % 3)          # strides
                hb = np.zeros(m * hs + 3, dt)
                hb[3::hs][:m] = h
                xb = np.zeros(n * xs, dt)
                xb[::xs] = x
                yb = np.zeros(lag * ys + 1, dt)
                y = fftm.correlate1d(ref, hb, View1(3, hs, m), xb, View1(0, xs, n), yb, View1(1, ys, lag),
                                     support, bias)
                got = y[1::ys][:lag]
                want = po.correlate1d(h.astype(dt), x.astype(dt), support, bias)
                err = po.rel_l2(got, want)
                assert err < tol, (m, n, support, bias, err)
