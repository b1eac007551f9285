"""GPU edge cases of the multiple FFT through the C ABI: the smallest
lengths and counts, the lengths the reference leaves undefined (defined
here), an out-of-place call whose input and output views are the same, and
vsip_fftm_getattr against the reference's own getattr on every plan type.
"""
import ctypes
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV, fftm_attr_f
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("major", [VSIP_ROW, VSIP_COL])
def test_length_one_is_identity_times_scale(amd, major):
    _case("ccfftmip_f length 1 (%s)" % ("ROW" if major == VSIP_ROW else "COL"))
    m, n = (7, 1) if major == VSIP_ROW else (1, 7)
    x = datagen.complex_data(140, (m, n))
    y = fftm.ccfftm(amd, x.ravel(), fftm.dense_view(m, n), major=major, scale=0.25)
    np.testing.assert_allclose(y, 0.25 * x.ravel(), rtol=1e-7, atol=0)
    assert not amd.last_error(), amd.last_error()


@pytest.mark.parametrize("m,n", [(1, 2), (1, 3), (5, 2), (3, 3), (1, 5), (1, 7), (2, 11), (1, 13), (1, 4096)])
def test_smallest_shapes_match_oracle_and_jvsip(amd, ref, m, n):
    _case("ccfftmip_f ROW %dx%d" % (m, n))
    x = datagen.complex_data(141, (m, n))
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v)
    assert po.rel_l2(y, po.ccfftm(x.ravel(), v)) < 1e-5
    assert po.rel_l2(y, fftm.ccfftm(ref, x.ravel(), v)) < 1e-5


def test_odd_real_length_is_refused(amd):
    _case("rcfftmop/crfftmop create with odd N returns NULL")
    import os
    os.environ["VSIP_AMD_QUIET"] = "1"
    assert not amd.rcfftmop_create_f(4, 1001, 1.0, VSIP_ROW, 0, 1)
    assert not amd.crfftmop_create_f(4, 1001, 1.0, VSIP_ROW, 0, 1)
    assert not amd.rcfftmop_create_f(3, 5, 1.0, VSIP_COL, 0, 1)
    amd.amd_clear_error()


@pytest.mark.parametrize("major", [VSIP_ROW, VSIP_COL])
def test_out_of_place_on_the_same_view_equals_in_place(amd, major):
    """ccfftmop(x, x): the reference copies x to y and transforms y in
    place; with y = x that is the in-place transform."""
    _case("ccfftmop_f with input view == output view (%s)" % ("ROW" if major == VSIP_ROW else "COL"))
    m, n = 12, 1000
    x = datagen.complex_data(142, (m, n))
    blk = np.ascontiguousarray(x).ravel().copy()
    b = amd.cblockbind_f(blk.ctypes.data, None, blk.size, 0)
    amd.cblockadmit_f(b, 1)
    v = amd.cmbind_f(b, 0, n, m, 1, n)
    p = amd.ccfftmop_create_f(m, n, 1.0, VSIP_FFT_FWD, major, 0, 1)
    amd.ccfftmop_f(p, v, v)
    amd.cblockrelease_f(b, 1, None, None)
    want = fftm.ccfftm(amd, x.ravel(), fftm.dense_view(m, n), major=major)
    assert po.rel_l2(blk, want) < 1e-6
    amd.fftm_destroy_f(p)
    amd.cmdestroy_f(v)
    amd.cblockdestroy_f(b)


def _attr(lib, p):
    a = fftm_attr_f()
    lib.fftm_getattr_f(p, ctypes.byref(a))
    return ((a.input.r, a.input.c), (a.output.r, a.output.c), a.place, a.scale, a.dir, a.major)


@pytest.mark.parametrize("major", [VSIP_ROW, VSIP_COL])
def test_getattr_matches_reference(amd, ref, major):
    _case("vsip_fftm_getattr_f on the four plan types (%s)" % ("ROW" if major == VSIP_ROW else "COL"))
    m, n = 6, 16
    for name, args in (("ccfftmip_create_f", (m, n, 0.5, VSIP_FFT_INV, major, 0, 1)),
                       ("ccfftmop_create_f", (m, n, 2.0, VSIP_FFT_FWD, major, 0, 1)),
                       ("rcfftmop_create_f", (m, n, 1.0, major, 0, 1)),
                       ("crfftmop_create_f", (m, n, 0.125, major, 0, 1))):
        pa, pr = getattr(amd, name)(*args), getattr(ref, name)(*args)
        assert pa and pr
        assert _attr(amd, pa) == _attr(ref, pr), name
        amd.fftm_destroy_f(pa)
        ref.fftm_destroy_f(pr)


def test_column_probe(amd):
    """vsip_amd_col_probe times the column transforms' access pattern and
    refuses shapes it cannot hold."""
    import ctypes
    g = ctypes.c_double(0.0)
    assert amd.amd_col_probe(1920, 1000, 2000, 1000, 4, 5, ctypes.byref(g)) == 0
    assert 100.0 < g.value < 8000.0, g.value
    print("  COL 1920 x 1000 pattern, 4 columns per workgroup: %.0f GB/s" % g.value, flush=True)
    for args in ((1920, 1000, 2000, 1000, 3), (4096, 1000, 2000, 1000, 8), (16, 100, 50, 100, 4),
                 (16, 100, 100, 99, 4), (0, 100, 100, 100, 4)):
        assert amd.amd_col_probe(*args, 5, ctypes.byref(g)) == -1, args
