"""Run VSIPL multiple FFTs on numpy buffers through any VsipLib.

Each call follows the VSIPL user-block sequence a C caller writes
(reference test c_VSIP_testing/ccfftmip_f.h:49-69 and the storage lifecycle
of VI_cblockbind_f_di.h / VI_cblockadmit_f_di.h / VI_cblockrelease_f_di.h):

    block = vsip_cblockbind_f(data, NULL, n, hint)     user memory
    vsip_cblockadmit_f(block, VSIP_TRUE)               -> HBM
    view  = vsip_cmbind_f(block, offset, col_stride, col_length, row_stride, row_length)
    plan  = vsip_ccfftmip_create_f(M, N, scale, dir, major, ntimes, hint)
    vsip_ccfftmip_f(plan, view)
    vsip_cblockrelease_f(block, VSIP_TRUE, &r, &i)     HBM -> user memory

so the same function drives libvsip_amd.so and a build of the reference.
A `View` is the stride descriptor (offset, col_stride, col_length,
row_stride, row_length) in elements of the block.
"""
import ctypes
from collections import namedtuple

import numpy as np

from .vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV, VSIP_MEM_NONE, VSIP_ALG_TIME, P

View = namedtuple("View", "offset col_stride col_length row_stride row_length")


def dense_view(m, n, major=VSIP_ROW):
    """The view vsip_cmcreate_f(m, n, major) makes (vsip_cmcreate_f.c:30-36)."""
    if major == VSIP_ROW:
        return View(0, n, m, 1, n)
    return View(0, 1, m, m, n)


def block_len(v):
    """Smallest block length that holds every element of view v."""
    hi = v.offset + max(0, (v.col_length - 1) * v.col_stride) + max(0, (v.row_length - 1) * v.row_stride)
    lo = v.offset + min(0, (v.col_length - 1) * v.col_stride) + min(0, (v.row_length - 1) * v.row_stride)
    assert lo >= 0, "view reaches below the block start"
    return hi + 1


def gather(blk, v):
    """Elements of view v of a 1-D block array as an (col_length, row_length) array."""
    i = np.arange(v.col_length)[:, None]
    j = np.arange(v.row_length)[None, :]
    return blk[v.offset + i * v.col_stride + j * v.row_stride]


def scatter(blk, v, a):
    i = np.arange(v.col_length)[:, None]
    j = np.arange(v.row_length)[None, :]
    blk[v.offset + i * v.col_stride + j * v.row_stride] = a


def _t(prec):
    return "f" if prec == 4 else "d"


class _CBlock:
    """A user complex block bound to a numpy buffer (interleaved or split)."""

    def __init__(self, lib, arr, split=False):
        self.lib, self.t = lib, _t(arr.real.dtype.itemsize)
        self.split = split
        if split:
            self.re = np.ascontiguousarray(arr.real).copy()
            self.im = np.ascontiguousarray(arr.imag).copy()
            ip = self.im.ctypes.data
            rp = self.re.ctypes.data
        else:
            self.buf = np.ascontiguousarray(arr).copy()
            rp, ip = self.buf.ctypes.data, None
        self.n = arr.size
        self.h = getattr(lib, "cblockbind_" + self.t)(rp, ip, self.n, VSIP_MEM_NONE)
        assert self.h, "cblockbind failed"
        rc = getattr(lib, "cblockadmit_" + self.t)(self.h, 1)
        assert rc == 0, "cblockadmit failed"

    def release(self):
        r, i = P(), P()
        getattr(self.lib, "cblockrelease_" + self.t)(self.h, 1, ctypes.byref(r), ctypes.byref(i))
        if self.split:
            return self.re + 1j * self.im
        return self.buf

    def destroy(self):
        getattr(self.lib, "cblockdestroy_" + self.t)(self.h)


class _RBlock:
    def __init__(self, lib, arr):
        self.lib, self.t = lib, _t(arr.dtype.itemsize)
        self.buf = np.ascontiguousarray(arr).copy()
        self.h = getattr(lib, "blockbind_" + self.t)(self.buf.ctypes.data, arr.size, VSIP_MEM_NONE)
        assert self.h, "blockbind failed"
        rc = getattr(lib, "blockadmit_" + self.t)(self.h, 1)
        assert rc == 0, "blockadmit failed"

    def release(self):
        getattr(self.lib, "blockrelease_" + self.t)(self.h, 1)
        return self.buf

    def destroy(self):
        getattr(self.lib, "blockdestroy_" + self.t)(self.h)


def ccfftm(lib, blk, view, major=VSIP_ROW, direction=VSIP_FFT_FWD, scale=1.0,
           out_blk=None, out_view=None, split=False):
    """In place (out_blk None) or out of place multiple complex FFT.

    blk: 1-D complex64/complex128 array (the block).  Returns the block(s)
    after release: a new array for in-place, (x_block, y_block) for out of place.
    split: both blocks split complex (True), or a pair (input, output).
    """
    t = _t(blk.real.dtype.itemsize)
    M, N = view.col_length, view.row_length
    split_in, split_out = (split, split) if isinstance(split, bool) else split
    b = _CBlock(lib, blk, split_in)
    v = getattr(lib, "cmbind_" + t)(b.h, *view)
    if out_blk is None:
        plan = getattr(lib, "ccfftmip_create_" + t)(M, N, scale, direction, major, 0, VSIP_ALG_TIME)
        assert plan, "ccfftmip_create failed"
        getattr(lib, "ccfftmip_" + t)(plan, v)
        res = b.release()
        getattr(lib, "fftm_destroy_" + t)(plan)
        getattr(lib, "cmdestroy_" + t)(v)
        b.destroy()
        return res
    ob = _CBlock(lib, out_blk, split_out)
    ov = getattr(lib, "cmbind_" + t)(ob.h, *out_view)
    plan = getattr(lib, "ccfftmop_create_" + t)(M, N, scale, direction, major, 0, VSIP_ALG_TIME)
    assert plan, "ccfftmop_create failed"
    getattr(lib, "ccfftmop_" + t)(plan, v, ov)
    res_x = b.release()
    res_y = ob.release()
    getattr(lib, "fftm_destroy_" + t)(plan)
    getattr(lib, "cmdestroy_" + t)(v)
    getattr(lib, "cmdestroy_" + t)(ov)
    b.destroy()
    ob.destroy()
    return res_x, res_y


def rcfftm(lib, xblk, xview, zblk, zview, major=VSIP_ROW, scale=1.0):
    """vsip_rcfftmop: real block/view -> complex block/view; returns the z block."""
    t = _t(xblk.dtype.itemsize)
    M, N = xview.col_length, xview.row_length
    xb = _RBlock(lib, xblk)
    zb = _CBlock(lib, zblk)
    xv = getattr(lib, "mbind_" + t)(xb.h, *xview)
    zv = getattr(lib, "cmbind_" + t)(zb.h, *zview)
    plan = getattr(lib, "rcfftmop_create_" + t)(M, N, scale, major, 0, VSIP_ALG_TIME)
    assert plan, "rcfftmop_create failed"
    getattr(lib, "rcfftmop_" + t)(plan, xv, zv)
    xb.release()
    res = zb.release()
    getattr(lib, "fftm_destroy_" + t)(plan)
    getattr(lib, "mdestroy_" + t)(xv)
    getattr(lib, "cmdestroy_" + t)(zv)
    xb.destroy()
    zb.destroy()
    return res


def crfftm(lib, zblk, zview, xblk, xview, major=VSIP_ROW, scale=1.0):
    """vsip_crfftmop: complex block/view -> real block/view; returns the x block."""
    t = _t(xblk.dtype.itemsize)
    M, N = xview.col_length, xview.row_length
    zb = _CBlock(lib, zblk)
    xb = _RBlock(lib, xblk)
    zv = getattr(lib, "cmbind_" + t)(zb.h, *zview)
    xv = getattr(lib, "mbind_" + t)(xb.h, *xview)
    plan = getattr(lib, "crfftmop_create_" + t)(M, N, scale, major, 0, VSIP_ALG_TIME)
    assert plan, "crfftmop_create failed"
    getattr(lib, "crfftmop_" + t)(plan, zv, xv)
    zb.release()
    res = xb.release()
    getattr(lib, "fftm_destroy_" + t)(plan)
    getattr(lib, "cmdestroy_" + t)(zv)
    getattr(lib, "mdestroy_" + t)(xv)
    zb.destroy()
    xb.destroy()
    return res


# ---- 1-D FFT (vsip_ccfftip / ccfftop / rcfftop / crfftop) -----------------

View1 = namedtuple("View1", "offset stride length")


def as_matrix(v):
    """The 1 x length matrix view of vector view v (for the oracle)."""
    return View(v.offset, v.length * v.stride, 1, v.stride, v.length)


def ccfft(lib, blk, view, direction=VSIP_FFT_FWD, scale=1.0, out_blk=None, out_view=None, split=False):
    """1-D complex FFT of vector view `view` of block `blk` (in place when
    out_blk is None); returns the block(s) after release like ccfftm."""
    t = _t(blk.real.dtype.itemsize)
    b = _CBlock(lib, blk, split)
    v = getattr(lib, "cvbind_" + t)(b.h, *view)
    if out_blk is None:
        plan = getattr(lib, "ccfftip_create_" + t)(view.length, scale, direction, 0, VSIP_ALG_TIME)
        assert plan, "ccfftip_create failed"
        getattr(lib, "ccfftip_" + t)(plan, v)
        res = b.release()
        getattr(lib, "fft_destroy_" + t)(plan)
        getattr(lib, "cvdestroy_" + t)(v)
        b.destroy()
        return res
    ob = _CBlock(lib, out_blk, split)
    ov = getattr(lib, "cvbind_" + t)(ob.h, *out_view)
    plan = getattr(lib, "ccfftop_create_" + t)(view.length, scale, direction, 0, VSIP_ALG_TIME)
    assert plan, "ccfftop_create failed"
    getattr(lib, "ccfftop_" + t)(plan, v, ov)
    res_x = b.release()
    res_y = ob.release()
    getattr(lib, "fft_destroy_" + t)(plan)
    getattr(lib, "cvdestroy_" + t)(v)
    getattr(lib, "cvdestroy_" + t)(ov)
    b.destroy()
    ob.destroy()
    return res_x, res_y


def rcfft(lib, xblk, xview, zblk, zview, scale=1.0):
    """vsip_rcfftop: real vector -> complex vector; returns the z block."""
    t = _t(xblk.dtype.itemsize)
    xb = _RBlock(lib, xblk)
    zb = _CBlock(lib, zblk)
    xv = getattr(lib, "vbind_" + t)(xb.h, *xview)
    zv = getattr(lib, "cvbind_" + t)(zb.h, *zview)
    plan = getattr(lib, "rcfftop_create_" + t)(xview.length, scale, 0, VSIP_ALG_TIME)
    assert plan, "rcfftop_create failed"
    getattr(lib, "rcfftop_" + t)(plan, xv, zv)
    xb.release()
    res = zb.release()
    getattr(lib, "fft_destroy_" + t)(plan)
    getattr(lib, "vdestroy_" + t)(xv)
    getattr(lib, "cvdestroy_" + t)(zv)
    xb.destroy()
    zb.destroy()
    return res


def crfft(lib, zblk, zview, xblk, xview, scale=1.0):
    """vsip_crfftop: complex vector -> real vector; returns the x block."""
    t = _t(xblk.dtype.itemsize)
    zb = _CBlock(lib, zblk)
    xb = _RBlock(lib, xblk)
    zv = getattr(lib, "cvbind_" + t)(zb.h, *zview)
    xv = getattr(lib, "vbind_" + t)(xb.h, *xview)
    plan = getattr(lib, "crfftop_create_" + t)(xview.length, scale, 0, VSIP_ALG_TIME)
    assert plan, "crfftop_create failed"
    getattr(lib, "crfftop_" + t)(plan, zv, xv)
    zb.release()
    res = xb.release()
    getattr(lib, "fft_destroy_" + t)(plan)
    getattr(lib, "cvdestroy_" + t)(zv)
    getattr(lib, "vdestroy_" + t)(xv)
    zb.destroy()
    xb.destroy()
    return res


def correlate1d(lib, h_blk, h_view, x_blk, x_view, y_blk, y_view, support=0, bias=0, split=False, ntimes=0):
    """1-D correlation of reference view h with data view x into view y
    (vsip_correlate1d for real blocks, vsip_ccorrelate1d for complex ones):
    the object is created for (h length, x length, support), each block is a
    user block admitted before and released after the call.  Returns the y
    block after release (split: complex blocks split)."""
    cplx = np.iscomplexobj(h_blk)
    t = _t(h_blk.real.dtype.itemsize)
    if cplx:
        blocks = [_CBlock(lib, a, split) for a in (h_blk, x_blk, y_blk)]
        bind, kind = "cvbind_", "ccorr1d"
    else:
        blocks = [_RBlock(lib, a) for a in (h_blk, x_blk, y_blk)]
        bind, kind = "vbind_", "corr1d"
    views = [getattr(lib, bind + t)(b.h, *v) for b, v in zip(blocks, (h_view, x_view, y_view))]
    obj = getattr(lib, kind + "_create_" + t)(h_view.length, x_view.length, support, ntimes, VSIP_ALG_TIME)
    assert obj, kind + "_create failed"
    getattr(lib, ("ccorrelate1d_" if cplx else "correlate1d_") + t)(obj, bias, *views)
    res = blocks[2].release()
    blocks[0].release()
    blocks[1].release()
    getattr(lib, kind + "_destroy_" + t)(obj)
    for v in views:
        getattr(lib, ("cvdestroy_" if cplx else "vdestroy_") + t)(v)
    for b in blocks:
        b.destroy()
    return res


def convolve1d(lib, h_blk, h_view, symm, x_blk, x_view, y_blk, y_view, D=1, support=0, ntimes=0):
    """Decimated 1-D convolution (vsip_convolve1d) of data view x with the
    kernel view h (expanded by symm) into view y, on real user blocks
    admitted before and released after; returns the y block."""
    t = _t(h_blk.dtype.itemsize)
    blocks = [_RBlock(lib, a) for a in (h_blk, x_blk, y_blk)]
    views = [getattr(lib, "vbind_" + t)(b.h, *v) for b, v in zip(blocks, (h_view, x_view, y_view))]
    obj = getattr(lib, "conv1d_create_" + t)(views[0], symm, x_view.length, D, support, ntimes, VSIP_ALG_TIME)
    assert obj, "conv1d_create failed"
    getattr(lib, "convolve1d_" + t)(obj, views[1], views[2])
    res = blocks[2].release()
    blocks[0].release()
    blocks[1].release()
    getattr(lib, "conv1d_destroy_" + t)(obj)
    for v in views:
        getattr(lib, "vdestroy_" + t)(v)
    for b in blocks:
        b.destroy()
    return res
