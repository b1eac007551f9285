"""Seeded random sweep of the multiple-FFT entry points against the oracle.

Each case draws, from its own seed, one transform of the VSIPL multiple-FFT
family (ccfftmip, ccfftmop, rcfftmop, crfftmop; FP32 or FP64; ROW or COL),
a transform length from one of the planner's classes (powers of two,
2^a 3^b 5^c 7^d, lengths with radix-11/13 stages, Bluestein primes,
lengths of the two- and three-pass splits, lengths 1..4), a vector count,
and for every view its own storage: row- or column-major in its block,
padded between vectors and offset from the block start.  So the cases
cross the planner's thresholds (power-of-two row kernels, compiled
schedules from 2^12 elements and the generic kernel below, splits,
Bluestein) with view geometries the hand-picked tests do not combine.

Every case goes through the C ABI with the VSIPL user-block sequence and
checks, as the reference's own tests do for one geometry
(c_VSIP_testing/ccfftmop_f.h:51-63):
  - the output view against the double-precision oracle, relative L2
    <= 1e-5 (FP32) or <= 1e-12 (FP64);
  - every block element outside the output view unchanged, bit for bit;
  - the input block of an out-of-place transform unchanged, bit for bit.
The oracle's cost is bounded per case (its prime stages are O(p) per
output), so the sweep runs in seconds.
"""
import math
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.fftm import View, View1
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV
from oracle import pyoracle as po

TOL = {4: 1e-5, 8: 1e-12}
NCASES = 128

_SMOOTH = sorted({2 ** a * 3 ** b * 5 ** c * 7 ** d
                  for a in range(17) for b in range(8) for c in range(7) for d in range(5)
                  if 5 <= 2 ** a * 3 ** b * 5 ** c * 7 ** d <= 65536
                  and 2 ** a * 3 ** b * 5 ** c * 7 ** d & (2 ** a * 3 ** b * 5 ** c * 7 ** d - 1)})
_ELEVEN = sorted({s * q for s in [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 20, 64, 96, 128, 256, 1024]
                  for q in (11, 13, 121, 143, 169)})
_PRIMES = [17, 19, 23, 29, 31, 37, 61, 97, 127, 251, 257, 509, 521, 1009, 2003, 4099, 8191]
_LARGE = [1 << 17, 1 << 18, 1 << 19, 1 << 20, 3 ** 9 * 4, 5 ** 6 * 8, 2 ** 10 * 3 ** 5, 10 ** 5,
          2 ** 12 * 7 ** 2, 40000 * 3]


def _largest_prime(n):
    p, big = 2, 1
    while p * p <= n:
        while n % p == 0:
            n //= p
            big = p
        p += 1
    return max(big, n) if n > 1 else big


def _draw(seed):
    """One case: (kind, prec, major, L, K, scale, direction, rng)."""
    rng = np.random.default_rng(1000 + seed)
    kind = ["ip", "ip", "op", "rc", "cr"][int(rng.integers(0, 5))]
    prec = 8 if rng.random() < 0.25 else 4
    major = VSIP_ROW if rng.random() < 0.55 else VSIP_COL
    cls = rng.random()
    if cls < 0.25:
        L = 1 << int(rng.integers(1, 17))
    elif cls < 0.50:
        L = int(rng.choice(_SMOOTH))
    elif cls < 0.62:
        L = int(rng.choice(_ELEVEN))
    elif cls < 0.80:
        L = int(rng.choice(_PRIMES)) * int(rng.choice([1, 1, 2, 4, 3]))
    elif cls < 0.93:
        L = int(rng.choice(_LARGE))
    else:
        L = int(rng.integers(1, 5))
    if kind in ("rc", "cr") and L % 2:
        L *= 2
    cap = (1 << 20) if prec == 4 else (1 << 19)
    cap = max(L, cap)
    p = _largest_prime(L)
    if p > 13:                                   # the oracle's prime stages cost L * p per vector
        cap = min(cap, max(L, int(3e8 // p)))
    kmax = max(1, cap // L)
    K = int(round(math.exp(rng.uniform(0.0, math.log(kmax))))) if kmax > 1 else 1
    K = max(1, min(K, kmax))
    scale = float(rng.choice([1.0, 0.5, 1.0 / max(L, 1)]))
    direction = VSIP_FFT_FWD if rng.random() < 0.5 else VSIP_FFT_INV
    return kind, prec, major, L, K, scale, direction, rng


def _layout(rng, m, n):
    """A view of m x n in a block of its own: row- or column-major storage,
    padded between vectors, sometimes offset."""
    off = int(rng.choice([0, 0, 0, 3, 64]))
    pad = int(rng.choice([0, 0, 1, 7, 32]))
    if rng.random() < 0.5:
        return View(off, n + pad, m, 1, n)
    return View(off, 1, m, m + pad, n)


def _outside(nblk, v):
    """Mask of the block elements the view does not cover."""
    mask = np.ones(nblk, bool)
    i = np.arange(v.col_length)[:, None]
    j = np.arange(v.row_length)[None, :]
    mask[(v.offset + i * v.col_stride + j * v.row_stride).ravel()] = False
    return mask


def _check(name, got, want, tol):
    err = po.rel_l2(got, want)
    print("  %s: rel L2 %.3e" % (name, err), flush=True)
    assert err < tol, "%s: rel L2 %.3e >= %.1e" % (name, err, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(NCASES))
def test_sweep(amd, seed):
    kind, prec, major, L, K, scale, direction, rng = _draw(seed)
    m, n = (K, L) if major == VSIP_ROW else (L, K)
    ctype = np.complex64 if prec == 4 else np.complex128
    rtype = np.float32 if prec == 4 else np.float64
    if kind in ("ip", "op"):
        xv = _layout(rng, m, n)
    elif kind == "rc":
        xv = _layout(rng, m, n)
        zv = _layout(rng, m, n // 2 + 1) if major == VSIP_ROW else _layout(rng, m // 2 + 1, n)
    else:
        zv = _layout(rng, m, n // 2 + 1) if major == VSIP_ROW else _layout(rng, m // 2 + 1, n)
        xv = _layout(rng, m, n)
    print("\n[case] seed %d: %s_%s %s %dx%d (length %d, %d vectors) scale %g dir %d views %s%s"
          % (seed, {"ip": "ccfftmip", "op": "ccfftmop", "rc": "rcfftmop", "cr": "crfftmop"}[kind],
             "f" if prec == 4 else "d", "ROW" if major == VSIP_ROW else "COL", m, n, L, K, scale,
             direction, tuple(xv), "" if kind in ("ip", "op") else " / " + str(tuple(zv))), flush=True)
    sys.stdout.flush()
    tol = TOL[prec]
    if kind == "ip":
        x = datagen.complex_data(seed, (fftm.block_len(xv),), dtype=ctype)
        y = fftm.ccfftm(amd, x.copy(), xv, major=major, direction=direction, scale=scale)
        want = po.ccfftm(x, xv, major=major, direction=direction, scale=scale)
        _check("vs oracle", y, want, tol)
        out = _outside(x.size, xv)
        np.testing.assert_array_equal(y[out], x[out])
    elif kind == "op":
        yv = _layout(rng, m, n)
        x = datagen.complex_data(seed, (fftm.block_len(xv),), dtype=ctype)
        y0 = datagen.complex_data(seed + 5000, (fftm.block_len(yv),), dtype=ctype)
        xr, y = fftm.ccfftm(amd, x.copy(), xv, major=major, direction=direction, scale=scale,
                            out_blk=y0.copy(), out_view=yv)
        np.testing.assert_array_equal(xr, x)
        want = po.ccfftm(x, xv, major=major, direction=direction, scale=scale, out_blk=y0, out_view=yv)
        _check("vs oracle", y, want, tol)
        out = _outside(y0.size, yv)
        np.testing.assert_array_equal(y[out], y0[out])
    elif kind == "rc":
        x = datagen.real_data(seed, (fftm.block_len(xv),), dtype=rtype)
        z0 = datagen.complex_data(seed + 5000, (fftm.block_len(zv),), dtype=ctype)
        z = fftm.rcfftm(amd, x.copy(), xv, z0.copy(), zv, major=major, scale=scale)
        _check("vs oracle", z, po.rcfftm(x, xv, z0, zv, major=major, scale=scale), tol)
        out = _outside(z0.size, zv)
        np.testing.assert_array_equal(z[out], z0[out])
    else:
        z = datagen.complex_data(seed, (fftm.block_len(zv),), dtype=ctype)   # not Hermitian
        x0 = datagen.real_data(seed + 5000, (fftm.block_len(xv),), dtype=rtype)
        x = fftm.crfftm(amd, z.copy(), zv, x0.copy(), xv, major=major, scale=scale)
        _check("vs oracle", x, po.crfftm(z, zv, x0, xv, major=major, scale=scale), tol)
        out = _outside(x0.size, xv)
        np.testing.assert_array_equal(x[out], x0[out])


def test_sweep_covers_the_planner_classes():
    """The drawn cases reach every length class, kind, precision and major."""
    seen = set()
    for s in range(NCASES):
        kind, prec, major, L, K, _, _, _ = _draw(s)
        seen.add(kind)
        seen.add(prec)
        seen.add(major)
        if L & (L - 1) == 0 and L >= 2:
            seen.add("pow2")
        if _largest_prime(L) > 13:
            seen.add("bluestein")
        if L * K >= (1 << 12):
            seen.add("compiled")
        if L * K < (1 << 12):
            seen.add("generic")
        if L > 65536:
            seen.add("split")
        if _largest_prime(L) in (11, 13):
            seen.add("radix11/13")
    for want in ("ip", "op", "rc", "cr", 4, 8, VSIP_ROW, VSIP_COL, "pow2", "bluestein", "compiled", "generic",
                 "split", "radix11/13"):
        assert want in seen, want


# ---- split complex storage ---------------------------------------------------

NCASES_SPLIT = 48


def _draw_split(seed):
    """A complex case on split blocks: (kind, prec, major, L, K, scale,
    direction, (split in, split out), rng), lengths from the same classes."""
    kind, prec, major, L, K, scale, direction, rng = _draw(20000 + seed)
    kind = "ip" if kind == "ip" else "op"
    split = (True, True) if kind == "ip" else [(True, True), (True, False), (False, True)][int(rng.integers(0, 3))]
    return kind, prec, major, L, K, scale, direction, split, rng


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(NCASES_SPLIT))
def test_sweep_split_storage(amd, seed):
    """ccfftmip / ccfftmop on split complex blocks (real and imaginary parts
    in two arrays; for out of place either side or both): against the
    oracle, against the same transform on interleaved blocks (1e-6 / 1e-13),
    every element outside the output view and the input block unchanged."""
    kind, prec, major, L, K, scale, direction, split, rng = _draw_split(seed)
    m, n = (K, L) if major == VSIP_ROW else (L, K)
    ctype = np.complex64 if prec == 4 else np.complex128
    tol, tol2 = TOL[prec], (1e-6 if prec == 4 else 1e-13)
    xv = _layout(rng, m, n)
    print("\n[case] split seed %d: ccfftm%s_%s %s %dx%d (length %d, %d vectors) scale %g dir %d split %s view %s"
          % (seed, kind, "f" if prec == 4 else "d", "ROW" if major == VSIP_ROW else "COL", m, n, L, K, scale,
             direction, split, tuple(xv)), flush=True)
    sys.stdout.flush()
    x = datagen.complex_data(seed + 300, (fftm.block_len(xv),), dtype=ctype)
    kw = dict(major=major, direction=direction, scale=scale)
    if kind == "ip":
        y = fftm.ccfftm(amd, x.copy(), xv, split=True, **kw)
        yi = fftm.ccfftm(amd, x.copy(), xv, **kw)
        want = po.ccfftm(x, xv, **kw)
        out = _outside(x.size, xv)
        np.testing.assert_array_equal(y[out], x[out])
    else:
        yv = _layout(rng, m, n)
        y0 = datagen.complex_data(seed + 7000, (fftm.block_len(yv),), dtype=ctype)
        xr, y = fftm.ccfftm(amd, x.copy(), xv, out_blk=y0.copy(), out_view=yv, split=split, **kw)
        np.testing.assert_array_equal(xr, x)
        yi = fftm.ccfftm(amd, x.copy(), xv, out_blk=y0.copy(), out_view=yv, **kw)[1]
        want = po.ccfftm(x, xv, out_blk=y0, out_view=yv, **kw)
        out = _outside(y0.size, yv)
        np.testing.assert_array_equal(y[out], y0[out])
    _check("split vs oracle", y, want, tol)
    _check("split vs interleaved", y, yi, tol2)


def test_split_sweep_covers_the_classes():
    """The split-storage cases reach both kinds, precisions and majors, each
    mix of split sides, compiled and generic sizes, the splits and Bluestein."""
    seen = set()
    for s in range(NCASES_SPLIT):
        kind, prec, major, L, K, _, _, split, _ = _draw_split(s)
        seen.update([kind, prec, major, split])
        seen.add("compiled" if L * K >= (1 << 12) else "generic")
        if L > 65536:
            seen.add("split")
        if _largest_prime(L) > 13:
            seen.add("bluestein")
    for want in ("ip", "op", 4, 8, VSIP_ROW, VSIP_COL, (True, True), (True, False), (False, True), "compiled",
                 "generic", "split", "bluestein"):
        assert want in seen, want


# ---- the 1-D family (vsip_ccfftip/op, rcfftop, crfftop) --------------------

NCASES_1D = 64
_PRIMES_1D = [17, 31, 97, 127, 257, 521, 1009, 4099, 8191]


def _draw_1d(seed):
    """One 1-D case: (kind, prec, length, scale, direction, rng)."""
    rng = np.random.default_rng(5000 + seed)
    kind = ["ip", "ip", "op", "rc", "cr"][int(rng.integers(0, 5))]
    prec = 8 if rng.random() < 0.25 else 4
    cls = rng.random()
    if cls < 0.28:
        L = 1 << int(rng.integers(1, 21))
    elif cls < 0.5:
        L = int(rng.choice(_SMOOTH)) * int(rng.choice([1, 1, 8, 27]))
    elif cls < 0.6:
        L = int(rng.choice(_ELEVEN))
    elif cls < 0.8:
        L = int(rng.choice(_PRIMES_1D)) * int(rng.choice([1, 2, 4]))
    elif cls < 0.88:
        L = int(rng.choice([1 << 21, 3 << 20, 3 ** 13]))        # the three-pass split
    else:
        L = int(rng.integers(1, 9))
    if kind in ("rc", "cr") and L % 2:
        L *= 2
    if prec == 8:
        L = min(L, 1 << 20)
    scale = float(rng.choice([1.0, 0.5, 1.0 / max(L, 1)]))
    direction = VSIP_FFT_FWD if rng.random() < 0.5 else VSIP_FFT_INV
    return kind, prec, L, scale, direction, rng


def _vlayout(rng, n):
    """A vector view of length n in a block of its own: stride +-1, 2 or 3,
    sometimes offset."""
    st = int(rng.choice([1, 1, 1, 2, 3, -1, -2]))
    off = int(rng.choice([0, 0, 5]))
    if st < 0:
        off += (n - 1) * -st
    return View1(off, st, n)


def _vlen(v):
    return v.offset + max(0, (v.length - 1) * v.stride) + 1


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(NCASES_1D))
def test_sweep_1d(amd, seed):
    kind, prec, L, scale, direction, rng = _draw_1d(seed)
    ctype = np.complex64 if prec == 4 else np.complex128
    rtype = np.float32 if prec == 4 else np.float64
    h = L // 2 + 1
    if kind in ("ip", "op"):
        xv = _vlayout(rng, L)
        yv = _vlayout(rng, L)
    elif kind == "rc":
        xv, yv = _vlayout(rng, L), _vlayout(rng, h)
    else:
        xv, yv = _vlayout(rng, h), _vlayout(rng, L)
    print("\n[case] 1-D seed %d: %s_%s length %d scale %g dir %d views %s -> %s"
          % (seed, {"ip": "ccfftip", "op": "ccfftop", "rc": "rcfftop", "cr": "crfftop"}[kind],
             "f" if prec == 4 else "d", L, scale, direction, tuple(xv), tuple(yv)), flush=True)
    sys.stdout.flush()
    tol = TOL[prec]
    xm, ym = fftm.as_matrix(xv), fftm.as_matrix(yv)
    if kind == "ip":
        x = datagen.complex_data(seed, (_vlen(xv),), dtype=ctype)
        y = fftm.ccfft(amd, x.copy(), xv, direction=direction, scale=scale)
        _check("vs oracle", y, po.ccfftm(x, xm, 0, direction, scale), tol)
        out = _outside(x.size, xm)
        np.testing.assert_array_equal(y[out], x[out])
    elif kind == "op":
        x = datagen.complex_data(seed, (_vlen(xv),), dtype=ctype)
        y0 = datagen.complex_data(seed + 5000, (_vlen(yv),), dtype=ctype)
        xr, y = fftm.ccfft(amd, x.copy(), xv, direction=direction, scale=scale, out_blk=y0.copy(), out_view=yv)
        np.testing.assert_array_equal(xr, x)
        _check("vs oracle", y, po.ccfftm(x, xm, 0, direction, scale, out_blk=y0, out_view=ym), tol)
        out = _outside(y0.size, ym)
        np.testing.assert_array_equal(y[out], y0[out])
    elif kind == "rc":
        x = datagen.real_data(seed, (_vlen(xv),), dtype=rtype)
        z0 = datagen.complex_data(seed + 5000, (_vlen(yv),), dtype=ctype)
        z = fftm.rcfft(amd, x.copy(), xv, z0.copy(), yv, scale=scale)
        _check("vs oracle", z, po.rcfftm(x, xm, z0, ym, 0, scale), tol)
        out = _outside(z0.size, ym)
        np.testing.assert_array_equal(z[out], z0[out])
    else:
        z = datagen.complex_data(seed, (_vlen(xv),), dtype=ctype)            # not Hermitian
        x0 = datagen.real_data(seed + 5000, (_vlen(yv),), dtype=rtype)
        x = fftm.crfft(amd, z.copy(), xv, x0.copy(), yv, scale=scale)
        _check("vs oracle", x, po.crfftm(z, xm, x0, ym, 0, scale), tol)
        out = _outside(x0.size, ym)
        np.testing.assert_array_equal(x[out], x0[out])


def test_sweep_1d_covers_the_classes():
    seen = set()
    for s in range(NCASES_1D):
        kind, prec, L, _, _, rng = _draw_1d(s)
        seen.update((kind, prec))
        if L & (L - 1) == 0 and L >= 2:
            seen.add("pow2")
        if _largest_prime(L) > 13:
            seen.add("bluestein")
        if L >= (1 << 17):
            seen.add("split")
        if L >= (1 << 21):
            seen.add("split3")
    for want in ("ip", "op", "rc", "cr", 4, 8, "pow2", "bluestein", "split", "split3"):
        assert want in seen, want
