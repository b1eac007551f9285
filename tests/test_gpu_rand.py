"""GPU parity of the VSIPL random fills (vsip_[c][vm]rand[un]_[fd]) through
the C ABI: the device fill must write the reference's values bit for bit
(its own C in oracle/_ref) in the reference's element order, leave the
generator in the reference's state (the next scalar draws agree), cross the
portable generator's marker skip inside a fill, and split fills into
several launches without changing a value.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from jvsip_amd.vsip import VSIP_ROW, VSIP_COL
from test_rand import A1, lcg_back, state

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _views(lib, t, kind, shape, layout):
    """(handle, reader, destroyer) of a view to fill"""
    dt = np.float32 if t == "f" else np.float64
    if kind in ("v", "cv"):
        n, stride = shape, layout
        cplx = kind == "cv"
        blk = getattr(lib, ("cblockcreate_" if cplx else "blockcreate_") + t)(n * abs(stride) + 3, 0)
        off = 1 if stride > 0 else 1 + (n - 1) * -stride
        v = getattr(lib, ("cvbind_" if cplx else "vbind_") + t)(blk, off, stride, n)

        def read():
            out = np.empty(n, np.complex64 if (cplx and t == "f") else np.complex128 if cplx else dt)
            if cplx:
                getattr(lib, "cvcopyto_user_" + t)(v, out.ctypes.data, None)
            else:
                getattr(lib, "vcopyto_user_" + t)(v, out.ctypes.data)
            return out

        def destroy():
            getattr(lib, ("cvalldestroy_" if cplx else "valldestroy_") + t)(v)
        return v, read, destroy
    m, n = shape
    cplx = kind == "cm"
    blk = getattr(lib, ("cblockcreate_" if cplx else "blockcreate_") + t)(m * n * 2 + 5, 0)
    if layout == "row":
        args = (2, 2 * n, m, 1, n)           # every other row of a 2n-wide parent
    elif layout == "col":
        args = (3, 1, m, 2 * m, n)           # columns 2m apart
    else:
        args = ((m - 1) * n + 1, -n, m, 1, n)   # rows reversed: row i at (m-1-i) n + 1
    v = getattr(lib, ("cmbind_" if cplx else "mbind_") + t)(blk, *args)

    def read():
        if cplx:
            out = np.empty((m, n), np.complex64 if t == "f" else np.complex128)
            getattr(lib, "cmcopyto_user_" + t)(v, VSIP_ROW, out.ctypes.data, None)
        else:
            out = np.empty((m, n), dt)
            getattr(lib, "mcopyto_user_" + t)(v, VSIP_ROW, out.ctypes.data)
        return out

    def destroy():
        getattr(lib, ("cmalldestroy_" if cplx else "malldestroy_") + t)(v)
    return v, read, destroy


CASES = []
for kind in ("v", "cv", "m", "cm"):
    for dist in ("u", "n"):
        for t in ("f", "d"):
            if kind in ("v", "cv"):
                for shape, layout in ((1000, 1), (777, 3), (500, -2)):
                    CASES.append((kind, dist, t, shape, layout))
            else:
                for layout in ("row", "col", "rev"):
                    CASES.append((kind, dist, t, (37, 53), layout))


@pytest.mark.parametrize("portable", [0, 1])
@pytest.mark.parametrize("kind,dist,t,shape,layout", CASES)
def test_fill_matches_reference(amd, ref, kind, dist, t, shape, layout, portable):
    fn = "%srand%s_%s" % (kind, dist, t)
    _case("%s %s %s %s" % (fn, shape, layout, "portable" if portable else "nonportable"))
    amd.amd_clear_error()
    got = []
    for lib in (amd, ref):
        h = lib.randcreate(42, 3, 2, 1 - portable)
        v, read, destroy = _views(lib, t, kind, shape, layout)
        getattr(lib, fn)(h, v)
        after = [lib.randu_f(h) for _ in range(3)] + [lib.crandn_d(h).r]
        got.append((read(), after))
        destroy()
        lib.randdestroy(h)
    np.testing.assert_array_equal(got[0][0], got[1][0])
    assert got[0][1] == got[1][1]
    assert not amd.last_error(), amd.last_error()


@pytest.mark.parametrize("fn,t", [("cmrandn", "f"), ("mrandu", "d"), ("vrandn", "f"), ("cvrandu", "d")])
def test_fill_crosses_marker_skip(amd, ref, fn, t):
    kind = fn[:fn.index("rand")]
    _case("%s_%s with X1 meeting X2 inside the fill" % (fn, t))
    amd.amd_clear_error()
    shape = 600 if kind in ("v", "cv") else (30, 41)
    for back in (1, 737, 2999):
        got = []
        for lib in (amd, ref):
            h = lib.randcreate(9, 1, 1, 0)
            s = state(h)
            s.X1 = lcg_back(A1, s.c1, s.X2, back)
            v, read, destroy = _views(lib, t, kind, shape, 1 if kind in ("v", "cv") else "col")
            getattr(lib, "%s_%s" % (fn, t))(h, v)
            got.append((read(), (s.X, s.X1, s.X2), lib.randn_f(h)))      # portable: every field is set
            destroy()
            lib.randdestroy(h)
        np.testing.assert_array_equal(got[0][0], got[1][0])
        assert got[0][1:] == got[1][1:], back


def test_fill_split_into_launches(amd, ref, monkeypatch):
    _case("cmrandn_f 64 x 300 in launches of at most 4000 draws")
    monkeypatch.setenv("VSIP_AMD_RNG_LAUNCH_DRAWS", "4000")
    amd.amd_clear_error()
    got = []
    for lib in (amd, ref):
        h = lib.randcreate(11, 1, 1, 0)
        s = state(h)
        s.X1 = lcg_back(A1, s.c1, s.X2, 50000)       # the skip falls in a later launch
        v, read, destroy = _views(lib, "f", "cm", (64, 300), "col")
        getattr(lib, "cmrandn_f")(h, v)
        got.append((read(), (s.X, s.X1, s.X2)))
        destroy()
        lib.randdestroy(h)
    np.testing.assert_array_equal(got[0][0], got[1][0])
    assert got[0][1] == got[1][1]
    assert not amd.last_error(), amd.last_error()


def test_large_fill_matches_reference(amd, ref):
    m, n = 1024, 2048
    _case("cmrandn_f %d x %d (12.6 M draws), nonportable and portable" % (m, n))
    for portable in (0, 1):
        got = []
        for lib in (amd, ref):
            h = lib.randcreate(2024, 1, 1, 1 - portable)
            v = lib.cmcreate_f(m, n, VSIP_COL, 0)
            lib.cmrandn_f(h, v)
            out = np.empty((m, n), np.complex64)
            lib.cmcopyto_user_f(v, VSIP_ROW, out.ctypes.data, None)
            got.append(out)
            lib.cmalldestroy_f(v)
            lib.randdestroy(h)
        np.testing.assert_array_equal(got[0], got[1])
