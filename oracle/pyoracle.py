"""Python side of the oracle (test infrastructure only).

Loads oracle/_build/liboracle.so (the double-precision restatement in
fftm_oracle.c) and, when it was built, oracle/_ref/libjvsip_ref.so (the
reference's own C compiled from its sources).  Only tests/, the smoke check
and bench.py's cpu_baseline leg use this module, and only as a checker.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libjvsip_ref.so")


class OView(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_long), ("col_stride", ctypes.c_long),
                ("col_length", ctypes.c_long), ("row_stride", ctypes.c_long),
                ("row_length", ctypes.c_long)]


_O = None


def lib():
    global _O
    if _O is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError("oracle not built: make -C oracle")
        d = ctypes.CDLL(ORACLE_SO, mode=ctypes.RTLD_LOCAL)
        P = ctypes.c_void_p
        d.oracle_dft_ld.argtypes = [ctypes.c_long, P, ctypes.c_int, P]
        d.oracle_dft_ld.restype = None
        d.oracle_fft_d.argtypes = [ctypes.c_long, P, ctypes.c_int, P]
        d.oracle_fft_d.restype = ctypes.c_int
        for name in ("oracle_ccfftm",):
            getattr(d, name).argtypes = [P, OView, P, OView, ctypes.c_int, ctypes.c_int, ctypes.c_double]
            getattr(d, name).restype = ctypes.c_int
        for name in ("oracle_rcfftm", "oracle_crfftm"):
            getattr(d, name).argtypes = [P, OView, P, OView, ctypes.c_int, ctypes.c_double]
            getattr(d, name).restype = ctypes.c_int
        _O = d
    return _O


def _ov(v):
    return OView(*[int(x) for x in v])


def dft_ld(x, sign=-1):
    """Brute-force long-double DFT of a 1-D complex vector."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    lib().oracle_dft_ld(x.size, x.ctypes.data, sign, y.ctypes.data)
    return y


def fft(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    assert lib().oracle_fft_d(x.size, x.ctypes.data, sign, y.ctypes.data) == 0
    return y


def ccfftm(blk, view, major=0, direction=-1, scale=1.0, out_blk=None, out_view=None):
    """Oracle ccfftmip (out_blk None) / ccfftmop; returns the output block (complex128)."""
    x = np.ascontiguousarray(blk, dtype=np.complex128)
    if out_blk is None:
        y, yv = x.copy(), view
    else:
        y, yv = np.ascontiguousarray(out_blk, dtype=np.complex128).copy(), out_view
    rc = lib().oracle_ccfftm(x.ctypes.data, _ov(view), y.ctypes.data, _ov(yv), major, direction, scale)
    assert rc == 0, rc
    return y


def rcfftm(xblk, xview, zblk, zview, major=0, scale=1.0):
    x = np.ascontiguousarray(xblk, dtype=np.float64)
    z = np.ascontiguousarray(zblk, dtype=np.complex128).copy()
    rc = lib().oracle_rcfftm(x.ctypes.data, _ov(xview), z.ctypes.data, _ov(zview), major, scale)
    assert rc == 0, rc
    return z


def crfftm(zblk, zview, xblk, xview, major=0, scale=1.0):
    z = np.ascontiguousarray(zblk, dtype=np.complex128)
    x = np.ascontiguousarray(xblk, dtype=np.float64).copy()
    rc = lib().oracle_crfftm(z.ctypes.data, _ov(zview), x.ctypes.data, _ov(xview), major, scale)
    assert rc == 0, rc
    return x


def hanning(n, prec="f"):
    """vsip_vcreate_hanning_f/_d (c_VSIP_src/vsip_vcreate_hanning_f.c:21-37,
    vsip_vcreate_hanning_d.c:21-37): w[k-1] = 0.5 (1 - cos(k t)), k = 1..n,
    t = 6.2831853071796 / (n + 1); the _f form rounds t and k t to float and
    takes the cosine in double."""
    k = np.arange(1, n + 1)
    if prec == "f":
        t = np.float32(6.2831853071796) / np.float32(n + 1)
        x = (t * k.astype(np.float32)).astype(np.float32)
        c = np.cos(x.astype(np.float64)).astype(np.float32)
        return (np.float32(0.5) * (np.float32(1.0) - c)).astype(np.float32)
    t = 6.2831853071796 / (n + 1)
    return 0.5 * (1.0 - np.cos(t * k.astype(np.float64)))


def komega(data, ts_taper, array_taper, gram, gain, rc_scale=1.0, cc_scale=-1.0, cc_dir=1):
    """One komega() step of the reference's k-omega beamformer
    (examples/komegaExamples/kw_c/kw.c:7-25) in double precision:
    taper every row by ts_taper and every column by array_taper
    (vsip_vmmul_f ROW, COL), rcfftmop ROW into an M x (N/2+1) COL-major
    matrix, ccfftmip COL with kw_init's arguments (kw.c:52 passes scale
    VSIP_FFT_FWD = -1 and dir 1, i.e. the inverse sign), then
    gram + gain * |.|^2 (mcmagsq, smmul, madd).  data: M x N real; returns the
    new gram (M x N/2+1, float64)."""
    data = np.asarray(data, dtype=np.float64)
    m, n = data.shape
    nf = n // 2 + 1
    x = data * np.asarray(ts_taper, np.float64)[None, :] * np.asarray(array_taper, np.float64)[:, None]
    z = np.zeros(m * nf, np.complex128)
    z = rcfftm(x.ravel(), (0, n, m, 1, n), z, (0, 1, m, m, nf), 0, rc_scale)
    y = ccfftm(z, (0, 1, m, m, nf), 1, cc_dir, cc_scale)
    y = y.reshape(nf, m).T                       # COL-major block -> (m, nf)
    return np.asarray(gram, np.float64) + gain * (y.real ** 2 + y.imag ** 2)


def corr_length(m, n):
    """FFT length of a correlation object: the smallest power of two above
    m + n - 1 (c_VSIP_src/vsip_ccorr1d_create_f.c:35-41)."""
    N = 2
    while N <= m + n - 1:
        N *= 2
    return N


def corr_divisors(m, n, support):
    """Divisors of the unbiased correlation at each output lag, as the
    reference's loops apply them (m <= n):
      FULL  VI_cvunbiasfull_f (vsip_ccorrelate1d_f.c:16-58): the first m lags
            by 1, 2, ..., m; then by m; the last min(m, n - 1) by their count
            down to 1
      SAME  VI_cvunbiassame_f (:60-99): i < m/2 by i + ceil(m/2); i < n - m/2
            by m; then by n + m/2 - i
      MIN   (:186-188): by m"""
    if support == 2:
        return np.full(n - m + 1, float(m))
    i = np.arange(n if support == 1 else m + n - 1, dtype=np.float64)
    if support == 1:
        n0 = m // 2
        return np.where(i < n0, i + n0 + (m & 1), np.where(i < n - n0, float(m), n + n0 - i))
    mn, l3 = m + n - 1, min(m, n - 1)
    return np.where(i < m, i + 1, np.where(i < mn - l3, float(m), mn - i))


def correlate1d(h, x, support=0, bias=0):
    """1-D correlation of reference h (length m) with data x (length n >= m),
    the steps of vsip_ccorrelate1d_f (c_VSIP_src/vsip_ccorrelate1d_f.c:123-196;
    the real form vsip_correlate1d_f.c:82-151 takes zero imaginary parts and
    keeps the real part) in double precision:
      xt = N - n zeros then x; ht = h then N - m zeros              (:135-153)
      z = FFT(conj(FFT(xt) conj(FFT(ht))) / N), forward transforms  (:155-161)
      y = z[N - mn + s : N - mn + s + lag_len], s = 0 (FULL),
          (m - 1)/2 (SAME), m - 1 (MIN)                             (:164-193)
      unbiased: y / corr_divisors(m, n, support)
    Returns complex128 for complex inputs, float64 for real ones."""
    real = not (np.iscomplexobj(h) or np.iscomplexobj(x))
    h = np.asarray(h, dtype=np.complex128)
    x = np.asarray(x, dtype=np.complex128)
    m, n = h.size, x.size
    mn, N = m + n - 1, corr_length(m, n)
    xt = np.zeros(N, np.complex128)
    xt[N - n:] = x
    ht = np.zeros(N, np.complex128)
    ht[:m] = h
    z = fft(np.conj(fft(xt, -1) * np.conj(fft(ht, -1))) / N, -1)
    s = 0 if support == 0 else (m - 1) // 2 if support == 1 else m - 1
    lag = mn if support == 0 else n if support == 1 else n - m + 1
    y = z[N - mn + s:N - mn + s + lag]
    if bias == 1:
        y = y / corr_divisors(m, n, support)
    return y.real.copy() if real else y.copy()


def conv_kernel(h, symm):
    """The kernel a convolution object holds (c_VSIP_src/vsip_conv1d_create_f.c:
    symm 0 h; 1 (VSIP_SYM_EVEN_LEN_ODD) h then h reversed without its last
    element, length 2 len(h) - 1; 2 (VSIP_SYM_EVEN_LEN_EVEN) h then h
    reversed, length 2 len(h))."""
    h = np.asarray(h, dtype=np.float64)
    if symm == 1:
        return np.concatenate([h, h[-2::-1]])
    if symm == 2:
        return np.concatenate([h, h[::-1]])
    return h.copy()


def conv_out_len(m, n, D, support):
    """Output length (vsip_conv1d_create_f.c, the switch on support) for
    kernel length m, data length n, decimation D."""
    if support == 0:
        return (m + n - 2) // D + 1
    if support == 1:
        return (n - 1) // D + 1
    return (n - 1) // D - (m - 1) // D + 1


def convolve1d(h, symm, x, D=1, support=0):
    """Decimated 1-D convolution as vsip_convolve1d_f computes it (its FFT
    methods, vsip_convolve1d_f.c, reduce to the linear convolution
    z = kernel * x): y[k] = z[k D + s], s = 0 (FULL), kernel length / 2
    (SAME), kernel length - 1 (MIN), in double precision."""
    k = conv_kernel(h, symm)
    x = np.asarray(x, dtype=np.float64)
    z = np.convolve(x, k)
    m, n = k.size, x.size
    s = 0 if support == 0 else m // 2 if support == 1 else m - 1
    L = conv_out_len(m, n, D, support)
    return z[s + D * np.arange(L)]


_REF = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """rrjudd/jvsip's C library built from its sources, bound like libvsip_amd.so."""
    global _REF
    if _REF is None:
        import sys
        sys.path.insert(0, os.path.dirname(HERE))
        from jvsip_amd.vsip import VsipLib
        _REF = VsipLib(REF_SO)
    return _REF


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
