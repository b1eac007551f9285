"""GPU parity for view geometries that the kernels cannot address from one
base with 32-bit offsets, through the C ABI.

VSIPL strides are signed (vsip_stride, vsip.h:129), so a view may walk its
block backwards.  Such views, and launches whose views span 2 GiB or more,
take host-side paths:
  - a view with a negative stride runs through a temporary staged plan
    (copy to contiguous scratch, transform, copy back);
  - a view spanning more than 2 GiB runs as several launches over slabs of
    vectors.
Both must give the oracle's result (FP32 rel L2 <= 1e-5).
"""
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL_F = 1e-5


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _check(name, got, want, tol):
    err = po.rel_l2(got, want)
    print("  %s: rel L2 %.3e" % (name, err), flush=True)
    assert err < tol, "%s: rel L2 %.3e >= %.1e" % (name, err, tol)


@pytest.mark.parametrize("major", [VSIP_ROW, VSIP_COL])
def test_negative_col_stride(amd, major):
    m, n = 12, 1000
    _case("ccfftmip_f %dx%d major %d, rows in reverse (col_stride -%d)" % (m, n, major, n))
    x = datagen.complex_data(40, (m * n,))
    v = fftm.View((m - 1) * n, -n, m, 1, n)
    y = fftm.ccfftm(amd, x, v, major=major, direction=VSIP_FFT_INV, scale=0.5)
    _check("vs oracle", y, po.ccfftm(x, v, major=major, direction=VSIP_FFT_INV, scale=0.5), TOL_F)


def test_negative_row_stride_out_of_place(amd):
    m, n = 7, 4096
    _case("ccfftmop_f ROW %dx%d, elements in reverse (row_stride -1) into a forward view" % (m, n))
    x = datagen.complex_data(41, (m * n,))
    xv = fftm.View(n - 1, n, m, -1, n)
    yv = fftm.dense_view(m, n)
    y0 = np.zeros(m * n, np.complex64)
    xr, y = fftm.ccfftm(amd, x, xv, out_blk=y0, out_view=yv)
    np.testing.assert_array_equal(xr, x)
    _check("vs oracle", y, po.ccfftm(x, xv, out_blk=y0, out_view=yv), TOL_F)


def test_negative_stride_real(amd):
    m, n = 5, 2048
    _case("rcfftmop_f ROW %dx%d from rows in reverse" % (m, n))
    x = datagen.real_data(42, (m * n,))
    xv = fftm.View((m - 1) * n, -n, m, 1, n)
    zv = fftm.dense_view(m, n // 2 + 1)
    z0 = np.zeros(m * (n // 2 + 1), np.complex64)
    z = fftm.rcfftm(amd, x, xv, z0, zv)
    _check("vs oracle", z, po.rcfftm(x, xv, z0, zv), TOL_F)


def test_view_spanning_more_than_2gib(amd):
    """Three 4096-point rows 2^27 + 4096 elements apart (a 2 GiB span of a
    device block): the transform runs in two launches."""
    m, n = 3, 4096
    stride = (1 << 27) + 4096
    total = (m - 1) * stride + n
    _case("ccfftmip_f ROW %dx%d, rows %d elements apart (%.2f GiB span)" % (m, n, stride, total * 8 / 2 ** 30))
    amd.amd_clear_error()                 # errors recorded by earlier tests
    b = amd.cblockcreate_f(total, 0)
    assert b, amd.last_error()
    x = datagen.complex_data(43, (m, n))
    rows = [amd.cmbind_f(b, i * stride, n, 1, 1, n) for i in range(m)]
    for i, rv in enumerate(rows):
        xi = np.ascontiguousarray(x[i])
        amd.cmcopyfrom_user_f(xi.ctypes.data, None, VSIP_ROW, rv)
    v = amd.cmbind_f(b, 0, stride, m, 1, n)
    p = amd.ccfftmip_create_f(m, n, 1.0, VSIP_FFT_FWD, VSIP_ROW, 0, 1)
    assert v and p, amd.last_error()
    amd.ccfftmip_f(p, v)
    y = np.empty((m, n), np.complex64)
    amd.cmcopyto_user_f(v, VSIP_ROW, y.ctypes.data, None)
    assert not amd.last_error(), amd.last_error()
    _check("vs numpy fp64", y, np.fft.fft(x.astype(np.complex128), axis=1), TOL_F)
    amd.fftm_destroy_f(p)
    for rv in rows:
        amd.cmdestroy_f(rv)
    amd.cmalldestroy_f(v)
