"""GPU parity of the multiple-FFT entry points against the oracle and jvsip.

Every case runs through the C ABI (libvsip_amd.so) with the VSIPL user-block
sequence bind -> admit -> cmbind -> create -> execute -> release, then compares
with the double-precision oracle (relative L2) and, where the reference
library was built, with jvsip's own output on the same input.

Tolerances: FP32 1e-5 relative L2 against jvsip and against the oracle;
FP64 1e-12 against the oracle and jvsip's _d path.
"""
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

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


@pytest.mark.parametrize("m,n", [(64, 1024), (4, 16), (3, 8), (5, 4096), (2, 2), (7, 32),
                                 (9, 64), (4, 256), (3, 512), (2, 2048), (2, 8192), (1, 16384)])
def test_ccfftmip_row_pow2(amd, ref, m, n):
    _case("ccfftmip_f ROW %dx%d fwd" % (m, n))
    x = datagen.complex_data(1, (m, n))
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v), TOL_F)
    _check("vs jvsip", y, fftm.ccfftm(ref, x.ravel(), v), TOL_F)


@pytest.mark.parametrize("n", [3, 5, 7, 11, 13, 6, 9, 10, 12, 14, 15, 24, 60, 96, 100, 143, 243, 1000, 1001,
                               1920, 2187, 3125, 6561, 1152, 4000])
def test_ccfftmip_row_mixed(amd, ref, n):
    m = 6
    _case("ccfftmip_f ROW %dx%d fwd" % (m, n))
    x = datagen.complex_data(2, (m, n))
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v), TOL_F)
    _check("vs jvsip", y, fftm.ccfftm(ref, x.ravel(), v), TOL_F)


@pytest.mark.parametrize("direction,scale", [(VSIP_FFT_INV, 1.0), (VSIP_FFT_INV, 1.0 / 1024),
                                             (VSIP_FFT_FWD, 0.5)])
def test_ccfftmip_dir_scale(amd, ref, direction, scale):
    _case("ccfftmip_f ROW 64x1024 dir %d scale %g" % (direction, scale))
    x = datagen.complex_data(11, (64, 1024))
    v = fftm.dense_view(64, 1024)
    y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=scale)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v, direction=direction, scale=scale), TOL_F)
    _check("vs jvsip", y, fftm.ccfftm(ref, x.ravel(), v, direction=direction, scale=scale), TOL_F)


def _np_rows(x, direction, scale):
    x = x.astype(np.complex128)
    n = x.shape[1]
    return (np.fft.fft(x, axis=1) if direction < 0 else np.fft.ifft(x, axis=1) * n) * scale


@pytest.mark.parametrize("m,dpp", [(4100, "0"), (4096 + 3 * 512 + 7, "0"), (2048 + 1, "0"),
                                   (4100, "1"), (4096 + 3 * 512 + 7, "1"), (2048 + 1, "1")])
def test_pipelined_4096_rows(amd, monkeypatch, m, dpp):
    """4096-point rows, at least 4 per workgroup of a grid of 2 per CU
    (256 CUs), take the pipelined persistent kernel (rows b, b + G, ...);
    the uneven row counts give some workgroups one row more, and the
    prefetch past the last row reads nothing.  Forward, inverse with scale,
    and out of place from rows 2N apart, against numpy's FP64 FFT.  dpp "1"
    selects the form with cross-lane (DPP) exchanges (VSIP_AMD_P2_DPP)."""
    monkeypatch.setenv("VSIP_AMD_P2_DPP", dpp)
    n = 4096
    x = datagen.complex_data(140, (m, n), dist="uniform")
    v = fftm.dense_view(m, n)
    for direction, scale in ((VSIP_FFT_FWD, 1.0), (VSIP_FFT_INV, 1.0 / n)):
        _case("ccfftmip_f ROW %dx%d dir %+d (pipelined rows)" % (m, n, direction))
        y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=scale).reshape(m, n)
        _check("vs numpy fp64", y, _np_rows(x, direction, scale), TOL_F)
    _case("ccfftmop_f ROW %dx%d from a 2N-wide parent (pipelined rows)" % (m, n))
    xp = np.zeros((m, 2 * n), np.complex64)
    xp[:, :n] = x
    xv = fftm.View(0, 2 * n, m, 1, n)
    xr, y = fftm.ccfftm(amd, xp.ravel(), xv, out_blk=np.zeros(m * n, np.complex64), out_view=v)
    np.testing.assert_array_equal(xr, xp.ravel())
    _check("vs numpy fp64", y.reshape(m, n), _np_rows(x, VSIP_FFT_FWD, 1.0), TOL_F)


@pytest.mark.parametrize("m,n,direction", [(65536 + 5, 256, VSIP_FFT_FWD), (32768 + 3, 512, VSIP_FFT_INV),
                                           (16384 + 1, 1024, VSIP_FFT_FWD), (16384 + 1, 1024, VSIP_FFT_INV),
                                           (8192 + 1, 2048, VSIP_FFT_FWD), (8192 + 3, 1024, VSIP_FFT_INV),
                                           (4096 + 1, 2048, VSIP_FFT_FWD)])
def test_pipelined_short_rows(amd, m, n, direction):
    """Rows of 256..2048 in groups of 4096 / n, at least 4 groups per
    workgroup of a grid of 2 per CU (256 CUs), take the pipelined kernel;
    the row counts leave a partial last group.  Inverse cases carry a
    scale.  Against numpy's FP64 FFT."""
    scale = 1.0 if direction == VSIP_FFT_FWD else 1.0 / n
    _case("ccfftmip_f ROW %dx%d dir %+d (pipelined groups)" % (m, n, direction))
    x = datagen.complex_data(142, (m, n), dist="uniform")
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v, direction=direction, scale=scale).reshape(m, n)
    _check("vs numpy fp64", y, _np_rows(x, direction, scale), TOL_F)


@pytest.mark.parametrize("m,direction", [(2048 + 5, VSIP_FFT_FWD), (2048 + 5, VSIP_FFT_INV),
                                         (4096 + 512 + 1, VSIP_FFT_FWD)])
def test_pipelined_8192_rows(amd, m, direction):
    """8192-point rows, at least 4 per workgroup of a grid of 2 per CU (256
    CUs), take the pipelined 8192-point kernel (stage-1 twiddles from LDS,
    the last stage's second butterfly twiddled by W_16^r times the first's);
    uneven rows per workgroup.  Against numpy's FP64 FFT."""
    n = 8192
    scale = 1.0 if direction == VSIP_FFT_FWD else 1.0 / n
    _case("ccfftmip_f ROW %dx%d dir %+d (pipelined 8192)" % (m, n, direction))
    x = datagen.complex_data(147, (m, n), dist="uniform")
    y = fftm.ccfftm(amd, x.ravel(), fftm.dense_view(m, n), direction=direction, scale=scale).reshape(m, n)
    _check("vs numpy fp64", y, _np_rows(x, direction, scale), TOL_F)


@pytest.mark.parametrize("m,direction", [(1024 + 3, VSIP_FFT_FWD), (1024 + 3, VSIP_FFT_INV),
                                         (2048 + 1, VSIP_FFT_FWD)])
def test_pipelined_16384_rows(amd, m, direction):
    """16384-point rows, at least 4 per workgroup of a grid of 1 per CU (256
    CUs), take the pipelined 16384-point kernel (radices 16 16 8 8, stage-1
    twiddles from LDS, the last stage's second butterfly twiddled by W_16^r
    times the first's); uneven rows per workgroup.  Against numpy's FP64
    FFT."""
    n = 16384
    scale = 1.0 if direction == VSIP_FFT_FWD else 1.0 / n
    _case("ccfftmip_f ROW %dx%d dir %+d (pipelined 16384)" % (m, n, direction))
    x = datagen.complex_data(148, (m, n), dist="uniform")
    y = fftm.ccfftm(amd, x.ravel(), fftm.dense_view(m, n), direction=direction, scale=scale).reshape(m, n)
    _check("vs numpy fp64", y, _np_rows(x, direction, scale), TOL_F)


@pytest.mark.parametrize("m,n", [(16384 + 3, 1024), (4096 + 7, 4096), (2048 + 3, 8192), (1024 + 5, 16384)])
def test_pipelined_rows_in_place_strided(amd, m, n):
    """ccfftmip on rows 2N apart in their parent (the left half of a 2N-wide
    matrix), in place, on the pipelined kernels; the right half must stay
    untouched.  Forward then inverse with scale, against numpy's FP64 FFT."""
    _case("ccfftmip_f ROW %dx%d in place, parent rows 2N (pipelined)" % (m, n))
    xp = datagen.complex_data(146, (m, 2 * n), dist="uniform")
    v = fftm.View(0, 2 * n, m, 1, n)
    y = fftm.ccfftm(amd, xp.ravel(), v).reshape(m, 2 * n)
    np.testing.assert_array_equal(y[:, n:], xp[:, n:])
    _check("fwd vs numpy fp64", y[:, :n], _np_rows(xp[:, :n], VSIP_FFT_FWD, 1.0), TOL_F)
    z = fftm.ccfftm(amd, y.ravel(), v, direction=VSIP_FFT_INV, scale=1.0 / n).reshape(m, 2 * n)
    np.testing.assert_array_equal(z[:, n:], xp[:, n:])
    _check("round trip", z[:, :n], xp[:, :n], 1e-6)


def test_fwd_inv_roundtrip(amd):
    _case("ccfftmip_f ROW 128x4096 fwd then inv 1/N")
    x = datagen.complex_data(2, (128, 4096), dist="uniform")
    v = fftm.dense_view(128, 4096)
    y = fftm.ccfftm(amd, x.ravel(), v)
    z = fftm.ccfftm(amd, y, v, direction=VSIP_FFT_INV, scale=1.0 / 4096)
    _check("round trip", z, x.ravel(), 1e-6)


@pytest.mark.parametrize("m,n", [(1920, 64), (64, 1920), (1000, 16), (16, 4096), (8, 3)])
def test_ccfftmip_col(amd, ref, m, n):
    _case("ccfftmip_f COL %dx%d fwd" % (m, n))
    x = datagen.complex_data(3, (m, n))
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v, major=VSIP_COL)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v, major=VSIP_COL), TOL_F)
    _check("vs jvsip", y, fftm.ccfftm(ref, x.ravel(), v, major=VSIP_COL), TOL_F)


def test_strided_view_kat_geometry(amd, ref):
    # the geometry of the reference's own test: cmbind(block, 4, 2, 8, 17, 4)
    _case("ccfftmip_f strided view offset 4 col_stride 2 row_stride 17")
    blk = datagen.complex_data(5, (1000,))
    v = fftm.View(4, 2, 8, 17, 4)
    for major in (VSIP_ROW, VSIP_COL):
        y = fftm.ccfftm(amd, blk, v, major=major)
        want = po.ccfftm(blk, v, major=major)
        _check("vs oracle major %d" % major, fftm.gather(y, v), fftm.gather(want, v), TOL_F)
        np.testing.assert_array_equal(np.delete(y, _idx(v)), np.delete(blk, _idx(v)))
        _check("vs jvsip major %d" % major, y, fftm.ccfftm(ref, blk, v, major=major), TOL_F)


def _idx(v):
    i = np.arange(v.col_length)[:, None]
    j = np.arange(v.row_length)[None, :]
    return (v.offset + i * v.col_stride + j * v.row_stride).ravel()


@pytest.mark.parametrize("major", [VSIP_ROW, VSIP_COL])
def test_ccfftmop_subview_parent2x(amd, ref, major):
    # rows lie 2N apart: the left half of a 2N-wide parent
    m, n = (64, 2187) if major == VSIP_ROW else (1920, 100)
    _case("ccfftmop_f major %d %dx%d parent row 2N" % (major, m, n))
    x = datagen.complex_data(3, (m * 2 * n,))
    xv = fftm.View(0, 2 * n, m, 1, n)
    y0 = np.zeros(m * n, np.complex64)
    yv = fftm.dense_view(m, n)
    xr, y = fftm.ccfftm(amd, x, xv, major=major, out_blk=y0, out_view=yv)
    np.testing.assert_array_equal(xr, x)  # input preserved
    _check("vs oracle", y, po.ccfftm(x, xv, major=major, out_blk=y0, out_view=yv), TOL_F)
    _, yr = fftm.ccfftm(ref, x, xv, major=major, out_blk=y0, out_view=yv)
    _check("vs jvsip", y, yr, TOL_F)


@pytest.mark.parametrize("m,n", [(16, 64), (8192 // 64, 2048), (3, 2), (5, 30), (4, 2000)])
def test_rcfftmop_row(amd, ref, m, n):
    _case("rcfftmop_f ROW %dx%d" % (m, n))
    x = datagen.real_data(2, (m, n))
    z0 = np.zeros(m * (n // 2 + 1), np.complex64)
    xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, n // 2 + 1)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv)
    _check("vs oracle", z, po.rcfftm(x.ravel(), xv, z0, zv), TOL_F)
    if n >= 4:  # jvsip divides by zero planning its n/2 = 1 point transform (VI_nuV.h)
        _check("vs jvsip", z, fftm.rcfftm(ref, x.ravel(), xv, z0, zv), TOL_F)


@pytest.mark.parametrize("m,n", [(16, 64), (128, 2048), (5, 30)])
def test_crfftmop_row(amd, ref, m, n):
    _case("crfftmop_f ROW %dx%d (non-Hermitian input)" % (m, n))
    z = datagen.complex_data(7, (m * (n // 2 + 1),))
    x0 = np.zeros(m * n, np.float32)
    zv, xv = fftm.dense_view(m, n // 2 + 1), fftm.dense_view(m, n)
    x = fftm.crfftm(amd, z, zv, x0, xv, scale=0.25)
    _check("vs oracle", x, po.crfftm(z, zv, x0, xv, scale=0.25), TOL_F)
    _check("vs jvsip", x, fftm.crfftm(ref, z, zv, x0, xv, scale=0.25), TOL_F)


def test_real_roundtrip_col(amd):
    m, n = 64, 48
    _case("rcfftmop_f/crfftmop_f COL %dx%d round trip" % (m, n))
    x = datagen.real_data(9, (m, n))
    xv = fftm.dense_view(m, n, VSIP_COL)
    zv = fftm.dense_view(m // 2 + 1, n, VSIP_COL)
    z = fftm.rcfftm(amd, x.ravel(), xv, np.zeros((m // 2 + 1) * n, np.complex64), zv, major=VSIP_COL)
    _check("rc vs oracle", z, po.rcfftm(x.ravel(), xv, np.zeros((m // 2 + 1) * n, np.complex64), zv, major=VSIP_COL), TOL_F)
    xb = fftm.crfftm(amd, z, zv, np.zeros(m * n, np.float32), xv, major=VSIP_COL, scale=1.0 / m)
    _check("round trip", xb, x.ravel(), 1e-6)


@pytest.mark.parametrize("m,n,major", [(8, 64, VSIP_ROW), (16, 2048, VSIP_ROW), (60, 40, VSIP_COL),
                                       (4, 1001, VSIP_ROW)])
def test_ccfftmip_d(amd, ref, m, n, major):
    _case("ccfftmip_d %dx%d major %d" % (m, n, major))
    x = datagen.complex_data(3, (m, n), dtype=np.complex128)
    v = fftm.dense_view(m, n)
    y = fftm.ccfftm(amd, x.ravel(), v, major=major)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v, major=major), TOL_D)
    _check("vs jvsip _d", y, fftm.ccfftm(ref, x.ravel(), v, major=major), TOL_D)


def test_split_complex(amd):
    _case("ccfftmip_f split block 8x64")
    x = datagen.complex_data(4, (8, 64))
    v = fftm.dense_view(8, 64)
    y = fftm.ccfftm(amd, x.ravel(), v, split=True)
    _check("vs oracle", y, po.ccfftm(x.ravel(), v), TOL_F)


@pytest.mark.parametrize("n", [65536, 32768, 40000])
def test_four_step(amd, n):
    m = 3
    _case("ccfftmip_f ROW %dx%d (two-pass)" % (m, n))
    x = datagen.complex_data(8, (m, n))
    v = fftm.dense_view(m, n)
    for d in (VSIP_FFT_FWD, VSIP_FFT_INV):
        y = fftm.ccfftm(amd, x.ravel(), v, direction=d, scale=0.5)
        _check("vs oracle dir %d" % d, y, po.ccfftm(x.ravel(), v, direction=d, scale=0.5), TOL_F)


@pytest.mark.parametrize("m,n", [(5, 512), (37, 1024), (3, 4096), (6, 8192), (2, 16384), (3, 32768)])
def test_real_pow2_rows(amd, ref, m, n):
    """rcfftmop/crfftmop on the power-of-two real-row kernel (half lengths
    256..16384), row counts that leave a partial last group."""
    _case("rcfftmop_f / crfftmop_f ROW %dx%d (power-of-two rows)" % (m, n))
    x = datagen.real_data(23, (m, n))
    xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, n // 2 + 1)
    z0 = np.zeros(m * (n // 2 + 1), np.complex64)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, scale=0.5)
    _check("rc vs oracle", z, po.rcfftm(x.ravel(), xv, z0, zv, scale=0.5), TOL_F)
    if n <= 4096:
        _check("rc vs jvsip", z, fftm.rcfftm(ref, x.ravel(), xv, z0, zv, scale=0.5), TOL_F)
    zz = datagen.complex_data(24, (m * (n // 2 + 1),))       # not Hermitian
    x0 = np.zeros(m * n, np.float32)
    xb = fftm.crfftm(amd, zz, zv, x0, xv, scale=1.0 / n)
    _check("cr vs oracle", xb, po.crfftm(zz, zv, x0, xv, scale=1.0 / n), TOL_F)
    if n <= 4096:
        _check("cr vs jvsip", xb, fftm.crfftm(ref, zz, zv, x0, xv, scale=1.0 / n), TOL_F)
    back = fftm.crfftm(amd, 2 * z, zv, x0, xv, scale=1.0 / n)
    _check("round trip", back, x.ravel(), 1e-6)


@pytest.mark.parametrize("m,parent", [(8195, 1), (8192 + 4 * 512 + 1, 2)])
def test_pipelined_real_rows(amd, m, parent):
    """rcfftmop on rows of 2048 reals with at least 4 groups of 4 rows per
    workgroup of a grid of 2 per CU (256 CUs) takes the pipelined R2C
    kernel; the row counts leave a partial last group and uneven groups per
    workgroup; parent 2 puts the complex rows 2(n/2+1) apart.  R2C against
    numpy's FP64 rfft; crfftmop (one group per workgroup) of a
    non-Hermitian input against the oracle, and the round trip."""
    n = 2048
    h = n // 2 + 1
    _case("rcfftmop_f ROW %dx%d, output parent x%d (pipelined R2C)" % (m, n, parent))
    x = datagen.real_data(141, (m, n))
    xv = fftm.dense_view(m, n)
    zv = fftm.View(0, parent * h, m, 1, h)
    z0 = np.zeros(m * parent * h, np.complex64)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, scale=0.25).reshape(m, parent * h)
    want = np.fft.rfft(x.astype(np.float64), axis=1) * 0.25
    _check("vs numpy fp64", z[:, :h], want, TOL_F)
    if parent > 1:
        assert not np.any(z[:, h:]), "stores outside the output rows"
    _case("crfftmop_f ROW %dx%d, input parent x%d" % (m, n, parent))
    zz = datagen.complex_data(143, (m * parent * h,))       # not Hermitian
    x0 = np.zeros(m * n, np.float32)
    xb = fftm.crfftm(amd, zz, zv, x0, xv, scale=1.0 / n)
    _check("vs oracle", xb, po.crfftm(zz, zv, x0, xv, scale=1.0 / n), TOL_F)
    back = fftm.crfftm(amd, (4 * z).ravel(), zv, x0, xv, scale=1.0 / n)
    _check("round trip", back, x.ravel(), 1e-6)


@pytest.mark.parametrize("m,n", [(32768 + 3, 512), (16384 + 5, 1024), (4096 + 1, 4096), (2048 + 1, 8192)])
def test_pipelined_r2c_lengths(amd, m, n):
    """The pipelined R2C kernel at the other half lengths (256, 512, 2048,
    4096: groups of 16, 8, 2, 1 rows), at least 4 groups per workgroup of a
    grid of 2 per CU (256 CUs), partial last group.  R2C against numpy's
    FP64 rfft; C2R (one group per workgroup) of a non-Hermitian input
    against the oracle, and the round trip."""
    h = n // 2 + 1
    _case("rcfftmop_f / crfftmop_f ROW %dx%d (pipelined R2C)" % (m, n))
    x = datagen.real_data(144, (m, n))
    xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, h)
    z = fftm.rcfftm(amd, x.ravel(), xv, np.zeros(m * h, np.complex64), zv, scale=2.0)
    _check("rc vs numpy fp64", z.reshape(m, h), np.fft.rfft(x.astype(np.float64), axis=1) * 2.0, TOL_F)
    x0 = np.zeros(m * n, np.float32)
    zz = datagen.complex_data(145, (m * h,))       # not Hermitian
    xb = fftm.crfftm(amd, zz, zv, x0, xv, scale=1.0 / n)
    _check("cr vs oracle", xb, po.crfftm(zz, zv, x0, xv, scale=1.0 / n), TOL_F)
    back = fftm.crfftm(amd, z, zv, x0, xv, scale=0.5 / n)
    _check("round trip", back, x.ravel(), 1e-6)


def test_real_pow2_strided_rows(amd):
    """Real rows 2x apart in their parents (rows of a 2N-wide real matrix,
    complex rows of a 2(N/2+1)-wide one) on the power-of-two real kernel."""
    m, n = 9, 2048
    h = n // 2 + 1
    _case("rcfftmop_f / crfftmop_f ROW %dx%d, parent rows 2x" % (m, n))
    x = datagen.real_data(25, (m * 2 * n,))
    xv = fftm.View(0, 2 * n, m, 1, n)
    zv = fftm.View(2, 2 * h, m, 1, h)
    z0 = np.zeros(m * 2 * h, np.complex64)
    z = fftm.rcfftm(amd, x, xv, z0, zv)
    _check("rc vs oracle", z, po.rcfftm(x, xv, z0, zv), TOL_F)
    x0 = np.zeros(m * 2 * n, np.float32)
    xb = fftm.crfftm(amd, z, zv, x0, xv, scale=1.0 / n)
    _check("cr vs oracle", xb, po.crfftm(z, zv, x0, xv, scale=1.0 / n), TOL_F)


# ---- FP64 real and out-of-place transforms (VI_rcfftmop_d_def.h, VI_crfftmop_d_def.h,
# VI_ccfftmop_d_def.h) against the oracle and the reference's _d at 1e-12 -------------

@pytest.mark.parametrize("m,n,major", [(16, 64, VSIP_ROW), (8, 2048, VSIP_ROW), (6, 1000, VSIP_ROW),
                                       (5, 30, VSIP_ROW), (3, 16384, VSIP_ROW), (60, 40, VSIP_COL),
                                       (32, 18, VSIP_COL), (2048, 3, VSIP_COL)])
def test_rcfftmop_d(amd, ref, m, n, major):
    _case("rcfftmop_d major %d %dx%d scale 0.5" % (major, m, n))
    x = datagen.real_data(31, (m, n), dtype=np.float64)
    if major == VSIP_ROW:
        xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, n // 2 + 1)
        nz = m * (n // 2 + 1)
    else:
        xv, zv = fftm.dense_view(m, n, VSIP_COL), fftm.dense_view(m // 2 + 1, n, VSIP_COL)
        nz = (m // 2 + 1) * n
    z0 = np.zeros(nz, np.complex128)
    z = fftm.rcfftm(amd, x.ravel(), xv, z0, zv, major=major, scale=0.5)
    _check("vs oracle", z, po.rcfftm(x.ravel(), xv, z0, zv, major=major, scale=0.5), TOL_D)
    _check("vs jvsip _d", z, fftm.rcfftm(ref, x.ravel(), xv, z0, zv, major=major, scale=0.5), TOL_D)


@pytest.mark.parametrize("m,n,major", [(16, 64, VSIP_ROW), (8, 2048, VSIP_ROW), (6, 1000, VSIP_ROW),
                                       (3, 16384, VSIP_ROW), (60, 40, VSIP_COL), (32, 18, VSIP_COL)])
def test_crfftmop_d(amd, ref, m, n, major):
    """non-Hermitian input: the literal pre-combine of VI_crfftmop_d_def.h"""
    _case("crfftmop_d major %d %dx%d (non-Hermitian input)" % (major, m, n))
    if major == VSIP_ROW:
        xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, n // 2 + 1)
        nz = m * (n // 2 + 1)
    else:
        xv, zv = fftm.dense_view(m, n, VSIP_COL), fftm.dense_view(m // 2 + 1, n, VSIP_COL)
        nz = (m // 2 + 1) * n
    z = datagen.complex_data(32, (nz,), dtype=np.complex128)
    x0 = np.zeros(m * n, np.float64)
    x = fftm.crfftm(amd, z, zv, x0, xv, major=major, scale=0.25)
    _check("vs oracle", x, po.crfftm(z, zv, x0, xv, major=major, scale=0.25), TOL_D)
    _check("vs jvsip _d", x, fftm.crfftm(ref, z, zv, x0, xv, major=major, scale=0.25), TOL_D)


def test_real_d_roundtrip(amd):
    m, n = 9, 4096
    _case("rcfftmop_d then crfftmop_d (scale 1/N) ROW %dx%d" % (m, n))
    x = datagen.real_data(34, (m, n), dtype=np.float64)
    xv, zv = fftm.dense_view(m, n), fftm.dense_view(m, n // 2 + 1)
    z = fftm.rcfftm(amd, x.ravel(), xv, np.zeros(m * (n // 2 + 1), np.complex128), zv)
    back = fftm.crfftm(amd, z, zv, np.zeros(m * n, np.float64), xv, scale=1.0 / n)
    _check("round trip", back, x.ravel(), 1e-14)


@pytest.mark.parametrize("major,m,n", [(VSIP_ROW, 64, 1001), (VSIP_ROW, 16, 2048), (VSIP_ROW, 8, 2187),
                                       (VSIP_COL, 300, 40), (VSIP_COL, 1920, 12)])
def test_ccfftmop_d_strided(amd, ref, major, m, n):
    """out of place from the left half of a 2N-wide parent, inverse with a scale"""
    _case("ccfftmop_d major %d %dx%d parent row 2N, inverse scale 1/3" % (major, m, n))
    x = datagen.complex_data(33, (m * 2 * n,), dtype=np.complex128)
    xv = fftm.View(0, 2 * n, m, 1, n)
    y0 = np.zeros(m * n, np.complex128)
    yv = fftm.dense_view(m, n)
    xr, y = fftm.ccfftm(amd, x, xv, major=major, direction=VSIP_FFT_INV, scale=1.0 / 3, out_blk=y0, out_view=yv)
    np.testing.assert_array_equal(xr, x)
    _check("vs oracle", y, po.ccfftm(x, xv, major=major, direction=VSIP_FFT_INV, scale=1.0 / 3,
                                     out_blk=y0, out_view=yv), TOL_D)
    _, yr = fftm.ccfftm(ref, x, xv, major=major, direction=VSIP_FFT_INV, scale=1.0 / 3, out_blk=y0, out_view=yv)
    _check("vs jvsip _d", y, yr, TOL_D)
