"""GPU tests of block storage semantics and the element operations, restating
the reference's own tests of this path through the C ABI:

  - c_VSIP_testing/cblock_admit_release_f.h: a split and an interleaved user
    block with the same values, admitted, scaled by 2 on the device
    (vsip_rscvmul_f), must read back equal element by element; released with
    update, the user arrays must hold the same values; rebound to swapped
    arrays (split <-> interleaved) and admitted again, the same checks hold;
  - c_VSIP_testing/ccfftm{ip,op}_{f,d}.h: the transform of the known-answer
    input, minus the expected output (vsip_cmsub), has a mean square value
    (vsip_cmmeansqval) within the test's bound -- computed here on the
    device with this library's own element operations;
  - the element operations themselves against numpy on strided views.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = ctypes.c_void_p


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def test_cblock_admit_release_rebind(amd):
    _case("cblock admit / rscvmul / release / rebind (cblock_admit_release_f.h)")
    amd.amd_clear_error()
    re_ = np.arange(10, dtype=np.float32)
    im_ = -np.arange(10, dtype=np.float32)
    ri = np.empty(20, np.float32)
    ri[0::2], ri[1::2] = re_, im_
    rb_re, rb_im, rb_ri = re_.copy(), im_.copy(), ri.copy()
    b_split = amd.cblockbind_f(re_.ctypes.data, im_.ctypes.data, 10, 0)
    b_il = amd.cblockbind_f(ri.ctypes.data, None, 10, 0)
    v_split = amd.cvbind_f(b_split, 0, 1, 10)
    v_il = amd.cvbind_f(b_il, 0, 1, 10)

    def admit_scale_compare():
        assert amd.cblockadmit_f(b_split, 1) == 0 and amd.cblockadmit_f(b_il, 1) == 0
        amd.rscvmul_f(2.0, v_split, v_split)
        amd.rscvmul_f(2.0, v_il, v_il)
        for i in range(10):
            a, b = amd.cvget_f(v_il, i), amd.cvget_f(v_split, i)
            assert (a.r, a.i) == (b.r, b.i), i

    def release_compare(first, second):
        p1r, p1i, p2r, p2i = P(), P(), P(), P()
        amd.cblockrelease_f(first, 1, ctypes.byref(p1r), ctypes.byref(p1i))
        amd.cblockrelease_f(second, 1, ctypes.byref(p2r), ctypes.byref(p2i))
        return p1r, p1i, p2r, p2i

    admit_scale_compare()
    s_r, s_i, i_r, i_i = release_compare(b_split, b_il)
    assert s_r.value == re_.ctypes.data and s_i.value == im_.ctypes.data
    assert i_r.value == ri.ctypes.data and i_i.value is None            # interleaved: NULL imaginary
    np.testing.assert_array_equal(ri[0::2], re_)
    np.testing.assert_array_equal(ri[1::2], im_)
    np.testing.assert_array_equal(re_, 2 * np.arange(10))

    # rebind: the interleaved block takes split arrays and vice versa; the
    # old user pointers come back
    o1r, o1i, o2r, o2i = P(), P(), P(), P()
    amd.cblockrebind_f(b_il, rb_re.ctypes.data, rb_im.ctypes.data, ctypes.byref(o1r), ctypes.byref(o1i))
    amd.cblockrebind_f(b_split, rb_ri.ctypes.data, None, ctypes.byref(o2r), ctypes.byref(o2i))
    assert o1r.value == ri.ctypes.data and o2r.value == re_.ctypes.data and o2i.value == im_.ctypes.data
    admit_scale_compare()
    release_compare(b_il, b_split)
    np.testing.assert_array_equal(rb_ri[0::2], rb_re)
    np.testing.assert_array_equal(rb_ri[1::2], rb_im)
    np.testing.assert_array_equal(rb_re, 2 * np.arange(10))
    amd.cvdestroy_f(v_split)
    amd.cvdestroy_f(v_il)
    amd.cblockdestroy_f(b_split)
    amd.cblockdestroy_f(b_il)
    assert not amd.last_error(), amd.last_error()


def _kats():
    with open(os.path.join(GOLDEN, "kat_fftm.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("k", range(len(_kats())))
def test_reference_kat_criterion_on_device(amd, k):
    """ccfftm{ip,op}_f.h: meansq(result - expected) on the device <= the
    test's bound."""
    c = _kats()[k]
    t = c["prec"]
    dt = np.complex64 if t == "f" else np.complex128
    _case("KAT %s %s major %d dir %+d via cmsub/cmmeansqval" % (c["source"], c["place"], c["major"], c["dir"]))
    amd.amd_clear_error()
    fn = lambda name: getattr(amd, name + "_" + t)
    x = (np.array(c["in_re"]) + 1j * np.array(c["in_im"])).astype(dt).reshape(c["M"], c["N"])
    want = (np.array(c["out_re"]) + 1j * np.array(c["out_im"])).astype(dt).reshape(c["M"], c["N"])
    blk = fn("cblockcreate")(c["block_len"], 0)
    xv = fn("cmbind")(blk, *fftm.View(*c["view"]))
    fn("cmcopyfrom_user")(np.ascontiguousarray(x).ctypes.data, None, VSIP_ROW, xv)
    if c["place"] == "ip":
        p = fn("ccfftmip_create")(c["M"], c["N"], c["scale"], c["dir"], c["major"], 0, 1)
        fn("ccfftmip")(p, xv)
        yv = xv
    else:
        yblk = fn("cblockcreate")(c["block_len"], 0)
        yv = fn("cmbind")(yblk, *fftm.View(*c["out_view"]))
        p = fn("ccfftmop_create")(c["M"], c["N"], c["scale"], c["dir"], c["major"], 0, 1)
        fn("ccfftmop")(p, xv, yv)
    ev = fn("cmcreate")(c["M"], c["N"], VSIP_ROW, 0)
    fn("cmcopyfrom_user")(np.ascontiguousarray(want).ctypes.data, None, VSIP_ROW, ev)
    dv = fn("cmcreate")(c["M"], c["N"], VSIP_ROW, 0)
    fn("cmsub")(yv, ev, dv)
    ms = fn("cmmeansqval")(dv)
    print("  meansq(result - expected) = %.3e (bound %.1e)" % (ms, c["tol_meansq"]), flush=True)
    assert ms <= c["tol_meansq"]
    assert not amd.last_error(), amd.last_error()
    fn("fftm_destroy")(p)
    for h in (ev, dv):
        fn("cmalldestroy")(h)
    if c["place"] != "ip":
        fn("cmalldestroy")(yv)
    fn("cmalldestroy")(xv)


def test_element_ops_strided(amd):
    _case("rscmmul / cmsub / cmmeansqval on strided views, f and d")
    amd.amd_clear_error()
    for t, dt in (("f", np.complex64), ("d", np.complex128)):
        m, n = 37, 53
        a = datagen.complex_data(90, (m, 2 * n), dtype=dt)
        b = datagen.complex_data(91, (m, 2 * n), dtype=dt)
        ba = getattr(amd, "cblockcreate_" + t)(m * 2 * n, 0)
        bb = getattr(amd, "cblockcreate_" + t)(m * 2 * n, 0)
        pa = getattr(amd, "cmbind_" + t)(ba, 0, 2 * n, m, 1, 2 * n)
        pb = getattr(amd, "cmbind_" + t)(bb, 0, 2 * n, m, 1, 2 * n)
        getattr(amd, "cmcopyfrom_user_" + t)(np.ascontiguousarray(a).ctypes.data, None, VSIP_ROW, pa)
        getattr(amd, "cmcopyfrom_user_" + t)(np.ascontiguousarray(b).ctypes.data, None, VSIP_ROW, pb)
        va = getattr(amd, "cmbind_" + t)(ba, 1, 2 * n, m, 2, n)            # odd columns
        vb = getattr(amd, "cmbind_" + t)(bb, 2 * n * (m - 1), -2 * n, m, 2, n)   # rows reversed
        out = getattr(amd, "cmcreate_" + t)(n, m, 1, 0)                    # COL-major m x n via transview
        ot = getattr(amd, "cmtransview_" + t)(out)
        getattr(amd, "cmsub_" + t)(va, vb, ot)
        got = np.empty((m, n), dt)
        getattr(amd, "cmcopyto_user_" + t)(ot, VSIP_ROW, got.ctypes.data, None)
        want = a[:, 1::2] - b[::-1, 0::2]
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-6 if t == "f" else 1e-14)
        ms = getattr(amd, "cmmeansqval_" + t)(ot)
        assert abs(ms - np.mean(np.abs(want) ** 2)) <= (1e-5 if t == "f" else 1e-12) * np.mean(np.abs(want) ** 2)
        getattr(amd, "rscmmul_" + t)(-0.5, va, va)
        getattr(amd, "cmcopyto_user_" + t)(va, VSIP_ROW, got.ctypes.data, None)
        np.testing.assert_allclose(got, -0.5 * a[:, 1::2], rtol=0, atol=1e-6 if t == "f" else 1e-14)
        for h in (va, vb, ot):
            getattr(amd, "cmdestroy_" + t)(h)
        getattr(amd, "cmalldestroy_" + t)(out)
        getattr(amd, "cmalldestroy_" + t)(pa)
        getattr(amd, "cmalldestroy_" + t)(pb)
    assert not amd.last_error(), amd.last_error()
