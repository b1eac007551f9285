"""GPU tests of 1-D correlation (vsip_[c]correlate1d_[fd]) through the C ABI.

Each case runs the VSIPL user-block sequence (bind, admit, bind views,
create, correlate, release) on libvsip_amd.so and checks y against the
oracle (pyoracle.correlate1d, pinned in test_corr_oracle.py): relative L2
<= 1e-5 (FP32) or 1e-12 (FP64), with every element of y's block outside the
view unchanged.  The reference's own build runs the same cases where it is
present.  Sizes reach N = 2^21 (data 2^20, reference 1024), where the
transforms run on the two- and three-pass splits.
"""
import sys

import numpy as np
import pytest

from jvsip_amd import fftm
from jvsip_amd.fftm import View1
from jvsip_amd.vsip import VSIP_SUPPORT_FULL, VSIP_SUPPORT_SAME, VSIP_SUPPORT_MIN, VSIP_BIASED, VSIP_UNBIASED
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SUPPORTS = (VSIP_SUPPORT_FULL, VSIP_SUPPORT_SAME, VSIP_SUPPORT_MIN)


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _lag(m, n, support):
    return m + n - 1 if support == 0 else n if support == 1 else n - m + 1


def _run(lib, h, x, support, bias, strides=(1, 1, 1), offsets=(0, 0, 0), split=False, seed=0):
    """Correlate on strided views of fresh blocks; returns (y, y block, y block before)."""
    m, n = h.size, x.size
    lag = _lag(m, n, support)
    dt = h.dtype
    hs, xs, ys = strides
    ho, xo, yo = offsets
    hb = np.zeros(ho + (m - 1) * abs(hs) + 1, dt)
    xb = np.zeros(xo + (n - 1) * abs(xs) + 1, dt)
    hv = View1(ho + (m - 1) * abs(hs) if hs < 0 else ho, hs, m)
    xv = View1(xo + (n - 1) * abs(xs) if xs < 0 else xo, xs, n)
    hb[hv.offset + hs * np.arange(m)] = h
    xb[xv.offset + xs * np.arange(n)] = x
    rng = np.random.default_rng(seed)
    y0 = (rng.standard_normal(yo + (lag - 1) * abs(ys) + 1 + 2) * (1 + 1j if np.iscomplexobj(h) else 1)).astype(dt)
    yv = View1(yo + (lag - 1) * abs(ys) if ys < 0 else yo, ys, lag)
    yb = fftm.correlate1d(lib, hb, hv, xb, xv, y0.copy(), yv, support, bias, split=split)
    return yb[yv.offset + ys * np.arange(lag)], yb, y0, yv


def _data(rng, m, n, cplx, prec):
    dt = (np.complex64 if prec == 4 else np.complex128) if cplx else (np.float32 if prec == 4 else np.float64)
    if cplx:
        return ((rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(dt),
                (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dt))
    return rng.standard_normal(m).astype(dt), rng.standard_normal(n).astype(dt)


SMALL = [(1, 1), (1, 9), (2, 2), (3, 3), (5, 12), (8, 8), (7, 20), (16, 100), (31, 64), (64, 65), (100, 1000),
         (257, 300), (1000, 1000)]


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("cplx", [True, False])
def test_correlate_against_oracle(amd, prec, cplx):
    """Every support and bias over lengths 1..1000 on unit-stride views."""
    rng = np.random.default_rng(100 + prec + cplx)
    tol = 1e-5 if prec == 4 else 1e-12
    for m, n in SMALL:
        h, x = _data(rng, m, n, cplx, prec)
        for support in SUPPORTS:
            for bias in (VSIP_BIASED, VSIP_UNBIASED):
                _case("%scorrelate1d_%s m %d n %d support %d bias %d" % ("c" if cplx else "", "f" if prec == 4 else "d",
                                                                         m, n, support, bias))
                y, yb, y0, yv = _run(amd, h, x, support, bias, seed=m + n)
                err = po.rel_l2(y, po.correlate1d(h, x, support, bias))
                assert err < tol, (m, n, support, bias, err)
                mask = np.ones(yb.size, bool)
                mask[yv.offset + yv.stride * np.arange(yv.length)] = False
                np.testing.assert_array_equal(yb[mask], y0[mask])
    assert not amd.last_error(), amd.last_error()


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("cplx", [True, False])
def test_correlate_strided_views(amd, prec, cplx):
    """Views with offsets and strides (negative ones included), and split
    complex blocks, against the oracle."""
    rng = np.random.default_rng(200 + prec + cplx)
    tol = 1e-5 if prec == 4 else 1e-12
    cases = [((2, 3, 1), (1, 0, 5)), ((-1, 2, 3), (0, 4, 1)), ((1, -2, -1), (3, 1, 0)), ((3, 1, 2), (0, 0, 0))]
    for (m, n) in [(7, 20), (64, 300), (129, 129)]:
        h, x = _data(rng, m, n, cplx, prec)
        for (strides, offsets) in cases:
            for support in SUPPORTS:
                for split in ((False, True) if cplx else (False,)):
                    _case("strided m %d n %d strides %s offsets %s support %d split %s"
                          % (m, n, strides, offsets, support, split))
                    y, _, _, _ = _run(amd, h, x, support, VSIP_UNBIASED, strides, offsets, split=split)
                    err = po.rel_l2(y, po.correlate1d(h, x, support, VSIP_UNBIASED))
                    assert err < tol, (m, n, strides, support, split, err)


@pytest.mark.parametrize("m,n", [(1024, 1 << 20), (4096, 300000), (1, 1 << 18)])
def test_correlate_long(amd, m, n):
    """Long data: N = 2^21 (two- and three-pass FFT splits), FP32 complex and
    real, against the oracle."""
    rng = np.random.default_rng(m + n)
    for cplx in (True, False):
        h, x = _data(rng, m, n, cplx, 4)
        for support in SUPPORTS:
            _case("long %scorrelate1d_f m %d n %d support %d" % ("c" if cplx else "", m, n, support))
            y, _, _, _ = _run(amd, h, x, support, VSIP_UNBIASED)
            err = po.rel_l2(y, po.correlate1d(h, x, support, VSIP_UNBIASED))
            assert err < 1e-5, (m, n, cplx, support, err)


@pytest.mark.parametrize("prec", [4, 8])
def test_correlate_matches_reference_build(amd, ref, prec):
    """The same inputs through rrjudd/jvsip's own build: FP32 within 1e-5,
    FP64 within 1e-12 of each other."""
    rng = np.random.default_rng(300 + prec)
    tol = 1e-5 if prec == 4 else 1e-12
    for cplx in (True, False):
        for m, n in [(5, 12), (64, 65), (100, 1000), (300, 4000)]:
            h, x = _data(rng, m, n, cplx, prec)
            for support in SUPPORTS:
                for bias in (VSIP_BIASED, VSIP_UNBIASED):
                    _case("vs reference build %s m %d n %d support %d bias %d" % (h.dtype, m, n, support, bias))
                    ya, _, _, _ = _run(amd, h, x, support, bias, (2, 1, 3), (1, 2, 0))
                    yr, _, _, _ = _run(ref, h, x, support, bias, (2, 1, 3), (1, 2, 0))
                    err = po.rel_l2(ya, yr)
                    assert err < tol, (cplx, m, n, support, bias, err)


def test_object_reuse_and_attributes(amd):
    """One object over two calls with other data (the transform buffer is
    reused), its attributes, and the refusals: M > N at create, views of the
    wrong length at the call."""
    import ctypes
    from jvsip_amd.vsip import corr1d_attr, VSIP_MEM_NONE, VSIP_ALG_TIME
    _case("ccorr1d object reuse and attributes")
    rng = np.random.default_rng(9)
    m, n = 33, 500
    obj = amd.ccorr1d_create_f(m, n, VSIP_SUPPORT_SAME, 0, VSIP_ALG_TIME)
    assert obj
    a = corr1d_attr()
    amd.ccorr1d_getattr_f(obj, ctypes.byref(a))
    assert (a.ref_len, a.data_len, a.support, a.lag_len) == (m, n, VSIP_SUPPORT_SAME, n)
    hv, xv, yv = amd.cvcreate_f(m, VSIP_MEM_NONE), amd.cvcreate_f(n, VSIP_MEM_NONE), amd.cvcreate_f(n, VSIP_MEM_NONE)
    for it in range(2):
        h, x = _data(rng, m, n, True, 4)
        parts = [np.ascontiguousarray(a) for a in (h.real, h.imag, x.real, x.imag)]    # alive across the calls
        amd.cvcopyfrom_user_f(parts[0].ctypes.data, parts[1].ctypes.data, hv)
        amd.cvcopyfrom_user_f(parts[2].ctypes.data, parts[3].ctypes.data, xv)
        amd.ccorrelate1d_f(obj, VSIP_BIASED, hv, xv, yv)
        yr, yi = np.empty(n, np.float32), np.empty(n, np.float32)
        amd.cvcopyto_user_f(yv, yr.ctypes.data, yi.ctypes.data)
        err = po.rel_l2(yr + 1j * yi, po.correlate1d(h, x, VSIP_SUPPORT_SAME, VSIP_BIASED))
        assert err < 1e-5, (it, err)
    assert not amd.last_error(), amd.last_error()
    amd.ccorrelate1d_f(obj, VSIP_BIASED, xv, xv, yv)            # h of length n: refused
    assert "length" in amd.last_error()
    amd.amd_clear_error()
    assert amd.ccorr1d_destroy_f(obj) == 0
    assert not amd.corr1d_create_f(10, 9, VSIP_SUPPORT_FULL, 0, VSIP_ALG_TIME)
    assert "M <= N" in amd.last_error()
    amd.amd_clear_error()
    for v in (hv, xv, yv):
        amd.cvalldestroy_f(v)


NSWEEP = 48


@pytest.mark.parametrize("seed", range(NSWEEP))
def test_correlate_sweep(amd, seed):
    """Seeded random correlation cases: lengths 1..3000 (m <= n), support,
    bias, real or complex, FP32 or FP64, strides -3..3 with offsets, split
    blocks for complex; against the oracle."""
    rng = np.random.default_rng(5000 + seed)
    cplx, prec = bool(rng.random() < 0.5), 8 if rng.random() < 0.3 else 4
    n = int(np.exp(rng.uniform(0, np.log(3000))))
    m = int(rng.integers(1, n + 1))
    support, bias = int(rng.integers(0, 3)), int(rng.integers(0, 2))
    strides = tuple(int(rng.choice([-3, -1, 1, 1, 2, 3])) for _ in range(3))
    offsets = tuple(int(rng.integers(0, 5)) for _ in range(3))
    split = cplx and bool(rng.random() < 0.3)
    _case("corr sweep %d: %s prec %d m %d n %d support %d bias %d strides %s offsets %s split %s"
          % (seed, "complex" if cplx else "real", prec, m, n, support, bias, strides, offsets, split))
    h, x = _data(rng, m, n, cplx, prec)
    y, _, _, _ = _run(amd, h, x, support, bias, strides, offsets, split=split)
    err = po.rel_l2(y, po.correlate1d(h, x, support, bias))
    assert err < (1e-5 if prec == 4 else 1e-12), err
