"""The convolution oracle (oracle/pyoracle.py: conv_kernel, conv_out_len,
convolve1d), pinned independently of numpy.convolve:

  - against a brute-force double loop of y[k] = sum_j h_j x_{kD + s - j}
    over every symmetry, support and decimation on small lengths;
  - the output lengths against the reference object's own attribute
    (vsip_conv1d_getattr_f: out_len, kernel_len) and the values against the
    reference's own build (oracle/_ref), FP32 <= 1e-5 and FP64 <= 1e-12,
    on every lag the reference computes: its sectioned method (data longer
    than twice the kernel, kernel shorter than 1000,
    c_VSIP_src/vsip_conv1d_create_f.c) loses the lags of z at and past
    (sections x section data length), which the mask below names;
  - a planted mistake (a SAME offset of (M - 1)/2, a kernel not reversed in
    the sum) fails.
"""
import ctypes

import numpy as np
import pytest

from jvsip_amd import fftm
from jvsip_amd.fftm import View1
from oracle import pyoracle as po


def _brute(k, x, D, support):
    m, n = k.size, x.size
    s = 0 if support == 0 else m // 2 if support == 1 else m - 1
    L = po.conv_out_len(m, n, D, support)
    y = np.zeros(L)
    for i in range(L):
        t = i * D + s
        for j in range(m):
            if 0 <= t - j < n:
                y[i] += k[j] * x[t - j]
    return y


CASES = [(1, 1), (1, 5), (2, 7), (3, 10), (4, 4), (5, 12), (8, 30), (13, 40)]


def test_matches_brute_force():
    rng = np.random.default_rng(1)
    for hl, n in CASES:
        for symm in (0, 1, 2):
            k = po.conv_kernel(rng.standard_normal(hl), symm)
            if k.size > n:
                continue
            x = rng.standard_normal(n)
            for D in (1, 2, 3):
                for support in (0, 1, 2):
                    if po.conv_out_len(k.size, n, D, support) < 1:
                        continue
                    h = k[:hl]
                    got = po.convolve1d(h, symm, x, D, support)
                    np.testing.assert_allclose(got, _brute(k, x, D, support), rtol=1e-12, atol=1e-12)


def test_kernel_expansion():
    h = np.array([1.0, 2.0, 3.0])
    np.testing.assert_array_equal(po.conv_kernel(h, 0), h)
    np.testing.assert_array_equal(po.conv_kernel(h, 1), [1, 2, 3, 2, 1])
    np.testing.assert_array_equal(po.conv_kernel(h, 2), [1, 2, 3, 3, 2, 1])


def test_planted_mistakes_fail():
    rng = np.random.default_rng(2)
    k, x = rng.standard_normal(6), rng.standard_normal(30)
    good = _brute(k, x, 1, 1)
    shifted = np.convolve(x, k)[(6 - 1) // 2:(6 - 1) // 2 + good.size]
    assert np.abs(shifted - good).max() > 1e-3
    unreversed = np.correlate(x, k, "full")[:good.size]
    assert np.abs(unreversed - good).max() > 1e-3


def _ref_good_lags(hl, symm, n, D, support, m):
    """Output positions the reference computes right: all but the lags of z
    at and past sections x data length in its sectioned method."""
    L = po.conv_out_len(m, n, D, support)
    keep = np.ones(L, bool)
    if n / hl > 2 and hl < 1000 and n / m > 2 and m < 1000:
        table = [(11, 32), (20, 64), (30, 128), (50, 256), (100, 512), (200, 1024), (300, 2048), (600, 4096)]
        nsec = next((v for lim, v in table if m < lim), 8192)
        ndata = nsec - m
        cut = -(-n // ndata) * ndata
        s = 0 if support == 0 else m // 2 if support == 1 else m - 1
        keep &= (s + D * np.arange(L)) < cut
    return keep


@pytest.mark.parametrize("prec", [4, 8])
def test_against_reference_build(ref, prec):
    from jvsip_amd.vsip import conv1d_attr
    rng = np.random.default_rng(40 + prec)
    dt = np.float32 if prec == 4 else np.float64
    tol = 1e-5 if prec == 4 else 1e-12
    t = "f" if prec == 4 else "d"
    for hl, n in [(1, 1), (3, 10), (4, 10), (5, 12), (8, 100), (16, 40), (7, 2000), (300, 1000), (1200, 1500),
                  (50, 5000), (513, 2000)]:
        for symm in (0, 1, 2):
            m = po.conv_kernel(np.zeros(hl), symm).size
            if m > n:
                continue
            for D in (1, 2, 3):
                for support in (0, 1, 2):
                    L = po.conv_out_len(m, n, D, support)
                    if L < 1:
                        continue
                    h, x = rng.standard_normal(hl).astype(dt), rng.standard_normal(n).astype(dt)
                    hb = np.ascontiguousarray(h)
                    blk = ref.__getattribute__("blockbind_" + t)(hb.ctypes.data, hl, 0)
                    ref.__getattribute__("blockadmit_" + t)(blk, 1)
                    hv = ref.__getattribute__("vbind_" + t)(blk, 0, 1, hl)
                    obj = ref.__getattribute__("conv1d_create_" + t)(hv, symm, n, D, support, 0, 1)
                    a = conv1d_attr()
                    ref.__getattribute__("conv1d_getattr_" + t)(obj, ctypes.byref(a))
                    assert (a.kernel_len, a.out_len, a.decimation) == (m, L, D)
                    ref.__getattribute__("conv1d_destroy_" + t)(obj)
                    ref.__getattribute__("vdestroy_" + t)(hv)
                    ref.__getattribute__("blockdestroy_" + t)(blk)
                    y = fftm.convolve1d(ref, h, View1(0, 1, hl), symm, x, View1(0, 1, n), np.zeros(L, dt),
                                        View1(0, 1, L), D, support)
                    want = po.convolve1d(h, symm, x, D, support)
                    keep = _ref_good_lags(hl, symm, n, D, support, m)
                    err = po.rel_l2(y[keep], want[keep]) if keep.any() else 0.0
                    assert err < tol, (hl, n, symm, D, support, err)
