"""GPU parity at the full sizes of the benchmark workloads, through the C ABI.

Each case is one of the workloads the benchmark times, at its real shape and
seed, checked the way the reference's own strided-view test checks a
transform (c_VSIP_testing/ccfftmop_f.h:51-63: bind a parent block, take a
subview, transform it out of place, compare):

  C2  vsip_ccfftmip_f forward, 4096 x 4096, U[-1,1] seed 2: the 8 rows of
      jvsip's output in tests/golden (<= 1e-5), the L2 norm of every row
      against jvsip's, the same rows against the oracle, and the inverse
      (scale 1/4096) back to the input.  rcfftmop_f 8192 x 2048 then
      crfftmop_f (scale 1/2048): sampled rows against jvsip's rcfftmop_f
      and the oracle, and the round trip.
  C3  vsip_ccfftmop_f on subviews whose rows lie 2x the row length apart:
      COL 1920 x 1000, ROW 1024 x 2187, ROW 512 x 1001, against jvsip's
      sampled rows and row norms (tests/golden) and the oracle; the input
      block (both halves of the parent) is left as it was.  The literal
      VSIPL reading (row_stride 2) at smaller sizes.  vsip_ccfftmip_d
      2048 x 2048 against jvsip's _d in full (<= 1e-12).
  C4  the 3000-job uneven batch (~2^27 elements): every job of the batched
      graph launch against the plain call of the same plan (<= 1e-6), a
      sample against the oracle (<= 1e-5), and three replays that agree bit
      for bit.
  C5  the 2000-request stream (rows U[1,512], lengths {256, 1000, 1024,
      2187, 4096, 16384}): served at 20k req/s, served with every request
      present at once (continuous batching under load), each request alone,
      and plain calls; every request of each form against the plain call on
      the device (<= 1e-6), a sample against the oracle (<= 1e-5).
"""
import os
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL_F = 1e-5
TOL_D = 1e-12


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _check(name, got, want, tol):
    err = po.rel_l2(got, want)
    print("  %s: rel L2 %.3e" % (name, err), flush=True)
    assert err < tol, "%s: rel L2 %.3e >= %.1e" % (name, err, tol)
    return err


def _norms_check(name, vecs, norms, tol):
    got = np.linalg.norm(vecs.astype(np.complex128), axis=1)
    err = float(np.max(np.abs(got - norms) / norms))
    print("  %s: max relative row-norm difference %.3e over %d rows" % (name, err, len(norms)), flush=True)
    assert err < tol, name


def _oracle_rows(vecs_in, direction=-1, scale=1.0):
    """the oracle's transform of each row of vecs_in"""
    m, n = vecs_in.shape
    return po.ccfftm(np.ascontiguousarray(vecs_in).ravel(), fftm.dense_view(m, n), direction=direction,
                     scale=scale).reshape(m, n)


def _dev_matrix(lib, x, t="f"):
    m, n = x.shape
    v = getattr(lib, "cmcreate_" + t)(m, n, VSIP_ROW, 0)
    assert v, lib.last_error()
    getattr(lib, "cmcopyfrom_user_" + t)(np.ascontiguousarray(x).ctypes.data, None, VSIP_ROW, v)
    return v


def _read(lib, v, shape, dtype=np.complex64, t="f"):
    y = np.empty(shape, dtype)
    getattr(lib, "cmcopyto_user_" + t)(v, VSIP_ROW, y.ctypes.data, None)
    return y


# ---- C2 ------------------------------------------------------------------------

def test_c2_4096_forward_inverse_vs_jvsip(amd):
    lib = amd
    f = np.load(os.path.join(GOLDEN, "jvsip_c2_4096x4096_fwd.npz"))
    m = n = 4096
    x = datagen.complex_data(2, (m, n), dist="uniform")
    _case("C2 ccfftmip_f ROW 4096x4096 forward (the headline kernel)")
    v = _dev_matrix(lib, x)
    fwd = lib.ccfftmip_create_f(m, n, 1.0, -1, VSIP_ROW, 0, 1)
    inv = lib.ccfftmip_create_f(m, n, 1.0 / n, 1, VSIP_ROW, 0, 1)
    assert fwd and inv, lib.last_error()
    lib.ccfftmip_f(fwd, v)
    y = _read(lib, v, (m, n))
    rows = f["rows"]
    _check("rows %s vs jvsip" % list(rows), y[rows], f["out_rows"], TOL_F)
    _norms_check("every row vs jvsip", y, f["row_norms"], TOL_F)
    _check("rows vs oracle", y[rows], _oracle_rows(x[rows]), TOL_F)
    _case("C2 ccfftmip_f ROW 4096x4096 inverse, scale 1/4096")
    lib.ccfftmip_f(inv, v)
    back = _read(lib, v, (m, n))
    _check("inverse(forward) vs input", back, x, 1e-6)
    assert not lib.last_error(), lib.last_error()
    lib.fftm_destroy_f(fwd)
    lib.fftm_destroy_f(inv)
    lib.cmalldestroy_f(v)


def test_c2_real_8192x2048_vs_jvsip(amd, ref):
    lib = amd
    m, n = 8192, 2048
    h = n // 2 + 1
    x = datagen.real_data(2, (m, n))
    _case("C2 rcfftmop_f ROW 8192x2048, then crfftmop_f scale 1/2048")
    xv = lib.mcreate_f(m, n, VSIP_ROW, 0)
    lib.mcopyfrom_user_f(np.ascontiguousarray(x).ctypes.data, VSIP_ROW, xv)
    zv = lib.cmcreate_f(m, h, VSIP_ROW, 0)
    bv = lib.mcreate_f(m, n, VSIP_ROW, 0)
    rc = lib.rcfftmop_create_f(m, n, 1.0, VSIP_ROW, 0, 1)
    cr = lib.crfftmop_create_f(m, n, 1.0 / n, VSIP_ROW, 0, 1)
    assert xv and zv and bv and rc and cr, lib.last_error()
    lib.rcfftmop_f(rc, xv, zv)
    lib.crfftmop_f(cr, zv, bv)
    z = _read(lib, zv, (m, h))
    back = np.empty((m, n), np.float32)
    lib.mcopyto_user_f(bv, VSIP_ROW, back.ctypes.data)
    rows = np.sort(np.random.default_rng(99).choice(m, 16, replace=False))
    sub = np.ascontiguousarray(x[rows])
    z0 = np.zeros(16 * h, np.complex64)
    want_j = fftm.rcfftm(ref, sub.ravel(), fftm.dense_view(16, n), z0, fftm.dense_view(16, h)).reshape(16, h)
    want_o = po.rcfftm(sub.ravel(), fftm.dense_view(16, n), z0, fftm.dense_view(16, h)).reshape(16, h)
    _check("16 rows vs jvsip rcfftmop_f", z[rows], want_j, TOL_F)
    _check("16 rows vs oracle", z[rows], want_o, TOL_F)
    _check("crfftmop(1/N)(rcfftmop(x)) vs x, every row", back, x, 1e-6)
    for p in (rc, cr):
        lib.fftm_destroy_f(p)
    lib.malldestroy_f(xv)
    lib.malldestroy_f(bv)
    lib.cmalldestroy_f(zv)


# ---- C3 ------------------------------------------------------------------------

C3 = [("jvsip_c3_col_1920x1000.npz", 1920, 1000, VSIP_COL, 2000),
      ("jvsip_c3_row_1024x2187.npz", 1024, 2187, VSIP_ROW, 4374),
      ("jvsip_c3_row_512x1001.npz", 512, 1001, VSIP_ROW, 2002)]


@pytest.mark.parametrize("fixture,m,n,major,parent", C3)
def test_c3_strided_ccfftmop_vs_jvsip(amd, fixture, m, n, major, parent):
    lib = amd
    f = np.load(os.path.join(GOLDEN, fixture))
    x = datagen.complex_data(3, (m * parent,))
    _case("C3 ccfftmop_f major %d %dx%d, rows of the %d-wide parent" % (major, m, n, parent))
    blk = lib.cblockcreate_f(m * parent, 0)
    pv = lib.cmbind_f(blk, 0, parent, m, 1, parent)
    lib.cmcopyfrom_user_f(x.ctypes.data, None, VSIP_ROW, pv)
    xv = lib.cmsubview_f(pv, 0, 0, m, n)
    yv = lib.cmcreate_f(m, n, VSIP_ROW, 0)
    p = lib.ccfftmop_create_f(m, n, 1.0, -1, major, 0, 1)
    assert blk and pv and xv and yv and p, lib.last_error()
    print("  schedule", lib.describe(p), flush=True)
    lib.ccfftmop_f(p, xv, yv)
    y = _read(lib, yv, (m, n))
    after = _read(lib, pv, (m * parent,))
    np.testing.assert_array_equal(after, x)            # x, and the parent's other half, untouched
    vecs = y if major == VSIP_ROW else y.T
    xs = fftm.gather(x, fftm.View(0, parent, m, 1, n))
    xin = xs if major == VSIP_ROW else xs.T
    sel = f["rows"]
    _check("vectors %s vs jvsip" % list(sel), vecs[sel], f["out_rows"], TOL_F)
    _norms_check("every vector vs jvsip", vecs, f["row_norms"], TOL_F)
    _check("vectors vs oracle", vecs[sel], _oracle_rows(xin[sel]), TOL_F)
    assert not lib.last_error(), lib.last_error()
    lib.fftm_destroy_f(p)
    lib.cmalldestroy_f(yv)
    lib.cmdestroy_f(xv)
    lib.cmalldestroy_f(pv)


@pytest.mark.parametrize("m,n,major", [(64, 2187, VSIP_ROW), (96, 1001, VSIP_ROW), (1920, 40, VSIP_COL)])
def test_c3_row_stride_2(amd, m, n, major):
    """The literal VSIPL reading of a 2x stride: elements of a row 2 apart
    (row_stride 2), rows 2n apart."""
    lib = amd
    x = datagen.complex_data(13, (m * 2 * n,))
    xv = fftm.View(0, 2 * n, m, 2, n)
    y0 = np.zeros(m * n, np.complex64)
    _case("ccfftmop_f major %d %dx%d, row_stride 2" % (major, m, n))
    xr, y = fftm.ccfftm(lib, x, xv, major=major, out_blk=y0, out_view=fftm.dense_view(m, n))
    np.testing.assert_array_equal(xr, x)
    _check("vs oracle", y, po.ccfftm(x, xv, major=major, out_blk=y0, out_view=fftm.dense_view(m, n)), TOL_F)


def test_c3_fp64_2048_vs_jvsip(amd, ref):
    lib = amd
    m = n = 2048
    x = datagen.complex_data(3, (m, n), dtype=np.complex128)
    _case("C3 ccfftmip_d ROW 2048x2048 forward, against jvsip's _d in full")
    v = _dev_matrix(lib, x, "d")
    p = lib.ccfftmip_create_d(m, n, 1.0, -1, VSIP_ROW, 0, 1)
    assert v and p, lib.last_error()
    lib.ccfftmip_d(p, v)
    y = _read(lib, v, (m, n), np.complex128, "d")
    want = fftm.ccfftm(ref, x.ravel(), fftm.dense_view(m, n)).reshape(m, n)
    _check("vs jvsip _d", y, want, TOL_D)
    rows = np.arange(0, m, 256)
    _check("rows vs oracle", y[rows], _oracle_rows(x[rows]), TOL_D)
    lib.fftm_destroy_d(p)
    lib.cmalldestroy_d(v)


# ---- C4 ------------------------------------------------------------------------

def test_c4_batch_3000_jobs(amd):
    lib = amd
    lengths, rows = datagen.c4_jobs(3000, 4)
    total = int(sum(r * n for r, n in zip(rows, lengths)))
    _case("C4 batch: %d jobs, %d distinct lengths, %d elements" % (len(lengths), len(set(lengths)), total))
    pool = datagen.complex_data(4, (total,))
    blk = lib.cblockcreate_f(total, 0)
    whole = lib.cmbind_f(blk, 0, total, 1, 1, total)
    rblk = lib.cblockcreate_f(total, 0)
    rwhole = lib.cmbind_f(rblk, 0, total, 1, 1, total)
    assert blk and whole and rblk and rwhole, lib.last_error()
    views, rviews, plans, off = [], [], [], 0
    for r, n in zip(rows, lengths):
        views.append(lib.cmbind_f(blk, off, n, r, 1, n))
        rviews.append(lib.cmbind_f(rblk, off, n, r, 1, n))
        # each plan runs a few times here (VSIPL's ntimes): no schedule compiles
        p = lib.ccfftmip_create_f(r, n, 1.0, -1, VSIP_ROW, 4, 1)
        assert p, lib.last_error()
        plans.append(p)
        off += r * n
    b = lib.batch(plans, views)
    assert b, lib.last_error()
    print("  batch", lib.batch_info(b), flush=True)
    lib.cmcopyfrom_user_f(pool.ctypes.data, None, VSIP_ROW, rwhole)
    for p, v in zip(plans, rviews):
        lib.ccfftmip_f(p, v)
    want = _read(lib, rwhole, (total,))
    got0 = None
    for rep in range(3):
        _case("C4 graph replay %d from the same input" % rep)
        lib.cmcopyfrom_user_f(pool.ctypes.data, None, VSIP_ROW, whole)
        assert lib.amd_batch_run_f(b) == 0, lib.last_error()
        got = _read(lib, whole, (total,))
        if got0 is None:
            got0 = got
        else:
            np.testing.assert_array_equal(got, got0)
    worst, off, sample = 0.0, 0, set(range(0, len(lengths), 125))
    for j, (r, n) in enumerate(zip(rows, lengths)):
        sl = slice(off, off + r * n)
        e = po.rel_l2(got0[sl], want[sl])
        worst = max(worst, e)
        assert e <= 1e-6, "job %d (%d x %d): batched vs plain %.3e" % (j, r, n, e)
        if j in sample:
            _check("job %d (%d x %d) vs oracle" % (j, r, n), got0[sl],
                   po.ccfftm(pool[sl], fftm.dense_view(r, n)), TOL_F)
        off += r * n
    print("  worst job, batched vs plain: %.3e" % worst, flush=True)
    assert not lib.last_error(), lib.last_error()
    lib.amd_batch_destroy_f(b)
    for p in plans:
        lib.fftm_destroy_f(p)
    for v in views + rviews:
        lib.cmdestroy_f(v)
    lib.cmalldestroy_f(whole)
    lib.cmalldestroy_f(rwhole)


# ---- C5 ------------------------------------------------------------------------

def test_c5_stream_2000_requests(amd):
    lib = amd
    reqs = datagen.c5_requests(2000, 20000.0, seed=5)
    offs = np.concatenate([[0], np.cumsum([r[1] * r[2] for r in reqs])])
    total = int(offs[-1])
    reqs = [r[:6] + (int(offs[i]),) for i, r in enumerate(reqs)]
    big = max(r[1] * r[2] for r in reqs)
    _case("C5 stream: %d requests, %d output elements (%.1f GB per form)" % (len(reqs), total, total * 8e-9))
    x = datagen.complex_data(5, (datagen.C5_POOL,))
    inb = lib.cblockcreate_f(datagen.C5_POOL, 0)
    inv = lib.cmbind_f(inb, 0, datagen.C5_POOL, 1, 1, datagen.C5_POOL)
    lib.cmcopyfrom_user_f(x.ctypes.data, None, VSIP_ROW, inv)
    forms = ["served", "saturated", "alone", "plain"]
    outs = {k: lib.cblockcreate_f(total, 0) for k in forms}
    assert all(outs.values()), lib.last_error()
    budget = 64 << 30
    _, st = lib.serve(reqs, inb, outs["served"], budget=budget, lanes=4, max_jobs=64)
    print("  served at 20k req/s:", st, flush=True)
    assert st["resident_requests"] > 0, "no request ran on the resident kernel"
    sat = [(0.0,) + r[1:] for r in reqs]
    _, st_sat = lib.serve(sat, inb, outs["saturated"], budget=budget, lanes=4, max_jobs=64)
    print("  served, all present at once:", st_sat, flush=True)
    assert st_sat["resident_requests"] + st_sat["batched_requests"] > 0, "no request ran in a batched launch"
    res_a, st_a = lib.serve(sat, inb, outs["alone"], budget=budget, lanes=1, max_jobs=1)
    assert np.all(res_a["t_admit"][1:] >= res_a["t_done"][:-1]), "alone: requests overlapped"
    _, st_p = lib.serve(reqs, inb, outs["plain"], budget=budget, plain=True)
    print("  plain calls:", st_p, flush=True)
    assert not lib.last_error(), lib.last_error()

    # every request of every form against the plain call, on the device
    dblk = lib.cblockcreate_f(big, 0)
    worst = dict.fromkeys(forms[:3], 0.0)
    for i, (t, r, n, d, s, io, oo) in enumerate(reqs):
        pv = lib.cmbind_f(outs["plain"], oo, n, r, 1, n)
        dv = lib.cmbind_f(dblk, 0, n, r, 1, n)
        ref_ms = lib.cmmeansqval_f(pv)
        for k in forms[:3]:
            v = lib.cmbind_f(outs[k], oo, n, r, 1, n)
            lib.cmsub_f(v, pv, dv)
            e = float(np.sqrt(lib.cmmeansqval_f(dv) / ref_ms))
            worst[k] = max(worst[k], e)
            assert e <= 1e-6, "request %d (%d x %d, dir %d): %s vs plain %.3e" % (i, r, n, d, k, e)
            lib.cmdestroy_f(v)
        lib.cmdestroy_f(pv)
        lib.cmdestroy_f(dv)
    print("  worst request vs plain:", worst, flush=True)
    for i in range(0, len(reqs), 125):
        t, r, n, d, s, io, oo = reqs[i]
        v = lib.cmbind_f(outs["served"], oo, n, r, 1, n)
        y = _read(lib, v, (r * n,))
        _check("request %d (%d x %d, dir %+d) vs oracle" % (i, r, n, d), y,
               po.ccfftm(x[io:io + r * n], fftm.dense_view(r, n), direction=d, scale=s), TOL_F)
        lib.cmdestroy_f(v)
    assert not lib.last_error(), lib.last_error()
    lib.cblockdestroy_f(dblk)
    for b in outs.values():
        lib.cblockdestroy_f(b)
    lib.cmalldestroy_f(inv)
