"""GPU tests of 1-D convolution (vsip_convolve1d_[fd]) through the C ABI:
every symmetry, decimation and support, FP32 and FP64, strided views and a
long case (kernel 1024, data 2^20), against the oracle (pyoracle.convolve1d,
pinned in test_conv_oracle.py; 1e-5 / 1e-12 relative L2), y's block outside
the view unchanged; against the reference's own build on the lags it
computes (its sectioned method loses the tail, see test_conv_oracle.py);
attributes and refusals."""
import ctypes
import sys

import numpy as np
import pytest

from jvsip_amd import fftm
from jvsip_amd.fftm import View1
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _run(lib, h, symm, x, D, support, strides=(1, 1, 1), offsets=(0, 0, 0), seed=0):
    hl, n = h.size, x.size
    m = po.conv_kernel(np.zeros(hl), symm).size
    L = po.conv_out_len(m, n, D, support)
    views, blocks = [], []
    for a, st, off in zip((h, x), strides[:2], offsets[:2]):
        b = np.zeros(off + (a.size - 1) * abs(st) + 1, a.dtype)
        v = View1(off + (a.size - 1) * abs(st) if st < 0 else off, st, a.size)
        b[v.offset + st * np.arange(a.size)] = a
        views.append(v)
        blocks.append(b)
    ys, yo = strides[2], offsets[2]
    y0 = np.random.default_rng(seed).standard_normal(yo + (L - 1) * abs(ys) + 3).astype(h.dtype)
    yv = View1(yo + (L - 1) * abs(ys) if ys < 0 else yo, ys, L)
    yb = fftm.convolve1d(lib, blocks[0], views[0], symm, blocks[1], views[1], y0.copy(), yv, D, support)
    return yb[yv.offset + ys * np.arange(L)], yb, y0, yv


@pytest.mark.parametrize("prec", [4, 8])
def test_convolve_against_oracle(amd, prec):
    rng = np.random.default_rng(10 + prec)
    dt = np.float32 if prec == 4 else np.float64
    tol = 1e-5 if prec == 4 else 1e-12
    for hl, n in [(1, 1), (1, 9), (3, 10), (5, 12), (8, 100), (16, 40), (7, 2000), (300, 1000), (1200, 1500),
                  (50, 5000)]:
        for symm in (0, 1, 2):
            m = po.conv_kernel(np.zeros(hl), symm).size
            for D in (1, 2, 3):
                for support in (0, 1, 2):
                    if support == 2 and (m > n or po.conv_out_len(m, n, D, support) < 1):
                        continue
                    _case("convolve1d_%s h %d symm %d n %d D %d support %d" % ("f" if prec == 4 else "d", hl, symm,
                                                                               n, D, support))
                    h, x = rng.standard_normal(hl).astype(dt), rng.standard_normal(n).astype(dt)
                    y, yb, y0, yv = _run(amd, h, symm, x, D, support, seed=hl + n)
                    err = po.rel_l2(y, po.convolve1d(h, symm, x, D, support))
                    assert err < tol, (hl, n, symm, D, support, err)
                    mask = np.ones(yb.size, bool)
                    mask[yv.offset + yv.stride * np.arange(yv.length)] = False
                    np.testing.assert_array_equal(yb[mask], y0[mask])
    assert not amd.last_error(), amd.last_error()


def test_convolve_strided_and_long(amd):
    rng = np.random.default_rng(3)
    for (hl, n, symm, D, support, strides, offsets) in [
            (9, 300, 1, 2, 1, (2, 3, 1), (1, 0, 5)), (9, 300, 0, 1, 0, (-1, 2, -3), (0, 4, 1)),
            (64, 1000, 2, 3, 2, (1, -1, 2), (3, 1, 0)), (1024, 1 << 20, 0, 1, 0, (1, 1, 1), (0, 0, 0)),
            (1024, 1 << 20, 1, 4, 1, (1, 1, 1), (0, 0, 0))]:
        _case("strided/long h %d n %d symm %d D %d support %d strides %s" % (hl, n, symm, D, support, strides))
        h, x = rng.standard_normal(hl).astype(np.float32), rng.standard_normal(n).astype(np.float32)
        y, _, _, _ = _run(amd, h, symm, x, D, support, strides, offsets)
        err = po.rel_l2(y, po.convolve1d(h, symm, x, D, support))
        assert err < 1e-5, (hl, n, symm, D, support, err)


@pytest.mark.parametrize("prec", [4, 8])
def test_convolve_matches_reference_build(amd, ref, prec):
    from test_conv_oracle import _ref_good_lags
    rng = np.random.default_rng(30 + prec)
    dt = np.float32 if prec == 4 else np.float64
    tol = 1e-5 if prec == 4 else 1e-12
    for hl, n in [(5, 12), (8, 100), (16, 40), (7, 2000), (300, 1000), (50, 5000)]:
        for symm in (0, 1, 2):
            m = po.conv_kernel(np.zeros(hl), symm).size
            for D in (1, 3):
                for support in (0, 1, 2):
                    if support == 2 and (m > n or po.conv_out_len(m, n, D, support) < 1):
                        continue
                    _case("vs reference build %s h %d symm %d n %d D %d support %d" % (dt.__name__, hl, symm, n, D,
                                                                                      support))
                    h, x = rng.standard_normal(hl).astype(dt), rng.standard_normal(n).astype(dt)
                    ya, _, _, _ = _run(amd, h, symm, x, D, support)
                    yr, _, _, _ = _run(ref, h, symm, x, D, support)
                    keep = _ref_good_lags(hl, symm, n, D, support, m)
                    err = po.rel_l2(ya[keep], yr[keep])
                    assert err < tol, (hl, n, symm, D, support, err)


def test_attributes_and_refusals(amd):
    from jvsip_amd.vsip import conv1d_attr, VSIP_MEM_NONE
    _case("conv1d attributes and refusals")
    h = np.arange(1, 6, dtype=np.float32)
    hv = amd.vcreate_f(5, VSIP_MEM_NONE)
    amd.vcopyfrom_user_f(h.ctypes.data, hv)
    obj = amd.conv1d_create_f(hv, 1, 100, 3, 1, 0, 1)
    assert obj
    a = conv1d_attr()
    amd.conv1d_getattr_f(obj, ctypes.byref(a))
    assert (a.kernel_len, a.symm, a.data_len, a.support, a.out_len, a.decimation) == (9, 1, 100, 1, 34, 3)
    assert amd.conv1d_destroy_f(obj) == 0
    assert not amd.conv1d_create_f(hv, 0, 3, 1, 2, 0, 1)          # MIN with the kernel longer than the data
    assert "MIN" in amd.last_error()
    amd.amd_clear_error()
    assert not amd.conv1d_create_f(hv, 0, 100, 0, 0, 0, 1)        # D = 0
    amd.amd_clear_error()
    amd.valldestroy_f(hv)


@pytest.mark.parametrize("seed", range(48))
def test_convolve_sweep(amd, seed):
    """Seeded random convolution cases: kernel 1..1500 (expanded by a random
    symmetry), data 1..5000, decimation 1..5, support, FP32 or FP64, strides
    -3..3 with offsets; against the oracle."""
    rng = np.random.default_rng(6000 + seed)
    prec = 8 if rng.random() < 0.3 else 4
    dt = np.float32 if prec == 4 else np.float64
    n = int(np.exp(rng.uniform(0, np.log(5000))))
    hl = int(np.exp(rng.uniform(0, np.log(1500))))
    symm, D, support = int(rng.integers(0, 3)), int(rng.integers(1, 6)), int(rng.integers(0, 3))
    m = po.conv_kernel(np.zeros(hl), symm).size
    if support == 2 and (m > n or po.conv_out_len(m, n, D, support) < 1):
        support = 0
    strides = tuple(int(rng.choice([-3, -1, 1, 1, 2, 3])) for _ in range(3))
    offsets = tuple(int(rng.integers(0, 5)) for _ in range(3))
    _case("conv sweep %d: prec %d h %d symm %d n %d D %d support %d strides %s offsets %s"
          % (seed, prec, hl, symm, n, D, support, strides, offsets))
    h, x = rng.standard_normal(hl).astype(dt), rng.standard_normal(n).astype(dt)
    y, _, _, _ = _run(amd, h, symm, x, D, support, strides, offsets)
    err = po.rel_l2(y, po.convolve1d(h, symm, x, D, support))
    assert err < (1e-5 if prec == 4 else 1e-12), err
