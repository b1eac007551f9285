"""GPU parity of the k-omega beamformer row (kw.c) through the C ABI.

  - the element operations komega() uses (vsip_vmmul_f ROW/COL, mcmagsq,
    smmul, madd, mfill, vcreate_hanning, _f and _d) against the reference's
    own C (oracle/_ref) bit for bit: both round every product and sum once;
  - real and imaginary part views (vsip_mrealview/mimagview) of interleaved
    and split blocks: get/put, copies and a real transform reading one;
  - komega() call for call on this library against the same calls on the
    reference's C (rel L2 <= 1e-5, FP32) and the FP64 oracle;
  - the fused step (vsip_amd_komega_f, two launches) against the VSIPL
    sequence and the oracle (rel L2 <= 1e-5), leaving its input unmodified.
"""
import ctypes
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, komega as kw
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _mat(lib, t, a, major=VSIP_ROW):
    a = np.ascontiguousarray(a)
    m = getattr(lib, "mcreate_" + t)(a.shape[0], a.shape[1], major, 0)
    getattr(lib, "mcopyfrom_user_" + t)(a.ctypes.data, VSIP_ROW, m)
    return m


def _get(lib, t, m, shape, dt):
    out = np.empty(shape, dt)
    getattr(lib, "mcopyto_user_" + t)(m, VSIP_ROW, out.ctypes.data)
    return out


def _vec(lib, t, a):
    a = np.ascontiguousarray(a)
    v = getattr(lib, "vcreate_" + t)(a.size, 0)
    getattr(lib, "vcopyfrom_user_" + t)(a.ctypes.data, v)
    return v


@pytest.mark.parametrize("t", ["f", "d"])
def test_real_element_ops_bit_exact_vs_reference(amd, ref, t):
    dt = np.float32 if t == "f" else np.float64
    m, n = 37, 61
    _case("vmmul ROW/COL, smmul, madd, mfill, mcmagsq, hanning (%s) vs jvsip's C" % t)
    amd.amd_clear_error()
    a = datagen.real_data(120, (m, n), dtype=dt)
    b = datagen.real_data(121, (m, n), dtype=dt)
    c = datagen.complex_data(122, (m, n), dtype=np.complex64 if t == "f" else np.complex128)
    vr, vc = datagen.real_data(123, (n,), dtype=dt), datagen.real_data(124, (m,), dtype=dt)
    outs = []
    for lib in (amd, ref):
        A = _mat(lib, t, a, VSIP_COL)              # COL storage: strides differ from the ROW result
        B = _mat(lib, t, b)
        R = getattr(lib, "mcreate_" + t)(m, n, VSIP_ROW, 0)
        C = getattr(lib, "cmcreate_" + t)(m, n, VSIP_COL, 0)
        getattr(lib, "cmcopyfrom_user_" + t)(np.ascontiguousarray(c).ctypes.data, None, VSIP_ROW, C)
        V1, V2 = _vec(lib, t, vr), _vec(lib, t, vc)
        res = {}
        getattr(lib, "vmmul_" + t)(V1, A, VSIP_ROW, R); res["vmmul_row"] = _get(lib, t, R, (m, n), dt)
        getattr(lib, "vmmul_" + t)(V2, A, VSIP_COL, R); res["vmmul_col"] = _get(lib, t, R, (m, n), dt)
        getattr(lib, "smmul_" + t)(0.3, A, R); res["smmul"] = _get(lib, t, R, (m, n), dt)
        getattr(lib, "madd_" + t)(A, B, R); res["madd"] = _get(lib, t, R, (m, n), dt)
        getattr(lib, "mcmagsq_" + t)(C, R); res["mcmagsq"] = _get(lib, t, R, (m, n), dt)
        getattr(lib, "vmmul_" + t)(V1, B, VSIP_ROW, B); res["vmmul_inplace"] = _get(lib, t, B, (m, n), dt)
        getattr(lib, "mfill_" + t)(-2.5, R); res["mfill"] = _get(lib, t, R, (m, n), dt)
        H = getattr(lib, "vcreate_hanning_" + t)(n, 0)
        h = np.empty(n, dt)
        getattr(lib, "vcopyto_user_" + t)(H, h.ctypes.data)
        res["hanning"] = h
        for x in (A, B, R):
            getattr(lib, "malldestroy_" + t)(x)
        getattr(lib, "cmalldestroy_" + t)(C)
        for x in (V1, V2, H):
            getattr(lib, "valldestroy_" + t)(x)
        outs.append(res)
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)
    np.testing.assert_array_equal(outs[0]["vmmul_row"], (vr[None, :] * a).astype(dt))
    np.testing.assert_array_equal(outs[0]["vmmul_col"], (vc[:, None] * a).astype(dt))
    assert not amd.last_error(), amd.last_error()


@pytest.mark.parametrize("split", [False, True])
def test_part_views(amd, split):
    m, n = 24, 40
    _case("mrealview/mimagview of a %s block" % ("split" if split else "interleaved"))
    amd.amd_clear_error()
    z = datagen.complex_data(125, (m, n))
    if split:
        re_, im_ = np.ascontiguousarray(z.real), np.ascontiguousarray(z.imag)
        blk = amd.cblockbind_f(re_.ctypes.data, im_.ctypes.data, m * n, 0)
    else:
        il = np.ascontiguousarray(z).view(np.float32).copy()
        blk = amd.cblockbind_f(il.ctypes.data, None, m * n, 0)
    assert amd.cblockadmit_f(blk, 1) == 0
    cv = amd.cmbind_f(blk, 0, n, m, 1, n)
    rv, iv = amd.mrealview_f(cv), amd.mimagview_f(cv)
    np.testing.assert_array_equal(_get(amd, "f", rv, (m, n), np.float32), z.real)
    np.testing.assert_array_equal(_get(amd, "f", iv, (m, n), np.float32), z.imag)
    assert amd.mget_f(rv, 3, 5) == z.real[3, 5] and amd.mget_f(iv, 7, 2) == z.imag[7, 2]
    amd.mput_f(iv, 1, 1, 9.0)
    # real part -> |.|^2 of the complex view, written into the real part
    amd.mcmagsq_f(cv, rv)
    want_re = np.abs(z) ** 2
    want_re[1, 1] = z.real[1, 1] ** 2 + 81.0
    np.testing.assert_allclose(_get(amd, "f", rv, (m, n), np.float32), want_re, rtol=1e-6)
    assert amd.mget_f(iv, 1, 1) == 9.0
    # a real transform reading the imaginary part (stride 2 scalars when interleaved)
    zb = np.zeros(m * (n // 2 + 1), np.complex64)
    zbk = amd.cblockbind_f(zb.ctypes.data, None, zb.size, 0)
    amd.cblockadmit_f(zbk, 1)
    zv = amd.cmbind_f(zbk, 0, n // 2 + 1, m, 1, n // 2 + 1)
    p = amd.rcfftmop_create_f(m, n, 1.0, VSIP_ROW, 0, 1)
    amd.rcfftmop_f(p, iv, zv)
    got = np.empty((m, n // 2 + 1), np.complex64)
    amd.cmcopyto_user_f(zv, VSIP_ROW, got.ctypes.data, None)
    im_now = z.imag.copy()
    im_now[1, 1] = 9.0
    want = np.fft.rfft(im_now.astype(np.float64), axis=1)
    assert po.rel_l2(got, want) < 1e-5
    for h in (rv, iv):
        amd.mdestroy_f(h)
    amd.fftm_destroy_f(p)
    amd.cmalldestroy_f(zv)
    amd.cmdestroy_f(cv)
    amd.cblockrelease_f(blk, 1, None, None)
    amd.cblockdestroy_f(blk)
    assert not amd.last_error(), amd.last_error()


KW_CASES = [(128, 1024, 4), (96, 1000, 2), (512, 4096, 2), (16, 64, 3)]


@pytest.mark.parametrize("m,n,navg", KW_CASES)
def test_komega_sequence_and_fused(amd, ref, m, n, navg):
    _case("k-omega %d sensors x %d samples, %d averages: VSIPL sequence, fused, jvsip, oracle" % (m, n, navg))
    amd.amd_clear_error()
    waves = [(n // 8, 3, 1.0), (n // 5, m - 7, 0.7), (n // 3, m // 2, 0.4)]
    seq, fus, jv = kw.KOmega(amd, m, n, navg), kw.KOmega(amd, m, n, navg), kw.KOmega(ref, m, n, navg)
    for k in (seq, fus, jv):
        k.zero()
    gram = np.zeros((m, n // 2 + 1))
    ts, arr = po.hanning(n, "f"), po.hanning(m, "f")
    for it in range(navg):
        x = kw.plane_waves(m, n, waves, seed=200 + it, noise=0.5)
        d1, d2, d3 = kw.data_matrix(amd, x), kw.data_matrix(amd, x), kw.data_matrix(ref, x)
        seq.step(d1)
        fus.step_fused(d2)
        jv.step(d3)
        np.testing.assert_array_equal(kw.read_matrix(amd, d2, (m, n)), x)      # fused: input untouched
        np.testing.assert_array_equal(kw.read_matrix(amd, d1, (m, n)), kw.read_matrix(ref, d3, (m, n)))
        for d, lib in ((d1, amd), (d2, amd), (d3, ref)):
            lib.malldestroy_f(d)
        gram = po.komega(x, ts, arr, gram, 1.0 / navg)
    g_seq, g_fus, g_jv = seq.gram(), fus.gram(), jv.gram()
    for k in (seq, fus, jv):
        k.destroy()
    for name, g, want, tol in (("sequence vs jvsip", g_seq, g_jv, 1e-5), ("sequence vs oracle", g_seq, gram, 1e-5),
                               ("fused vs sequence", g_fus, g_seq, 1e-5), ("fused vs oracle", g_fus, gram, 1e-5)):
        err = po.rel_l2(g, want)
        print("  %s: rel L2 %.3e" % (name, err), flush=True)
        assert err < tol, name
    i, j = np.unravel_index(np.argmax(g_fus), g_fus.shape)
    assert (j, i) == (n // 8, (-3) % m)
    assert not amd.last_error(), amd.last_error()


def test_komega_fused_rejects_other_layouts(amd):
    _case("vsip_amd_komega_f argument checks")
    amd.amd_clear_error()
    import os
    os.environ["VSIP_AMD_QUIET"] = "1"
    k = kw.KOmega(amd, 16, 64, 1)
    d = kw.data_matrix(amd, np.zeros((16, 64), np.float32))
    # a gram with other strides than freq
    g_row = amd.mcreate_f(16, 33, VSIP_ROW, 0)
    assert amd.amd_komega_f(k.rcfftm, k.ccfftm, k.ts_taper, k.array_taper, d, k.cm_freq, g_row, 1.0) == -1
    assert "strides" in amd.last_error()
    amd.amd_clear_error()
    # plans swapped
    assert amd.amd_komega_f(k.ccfftm, k.rcfftm, k.ts_taper, k.array_taper, d, k.cm_freq, k.m_gram, 1.0) == -1
    amd.amd_clear_error()
    # taper of the wrong length
    assert amd.amd_komega_f(k.rcfftm, k.ccfftm, k.array_taper, k.array_taper, d, k.cm_freq, k.m_gram, 1.0) == -1
    amd.amd_clear_error()
    amd.malldestroy_f(g_row)
    amd.malldestroy_f(d)
    k.destroy()
