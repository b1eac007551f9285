"""The k-omega beamformer of rrjudd/jvsip's example
(examples/komegaExamples/kw_c/kw.c) over the VSIPL C API.

`KOmega(lib, nsens, nts, navg)` builds kw_init's objects (kw.c:28-66) on any
library that exports the API -- libvsip_amd.so or the reference's own C
build -- and `step(data)` is komega() (kw.c:7-25) call for call.  On
libvsip_amd.so, `step_fused(data)` runs the same step as the library's two
fused launches (vsip_amd_komega_f).
"""
import ctypes

import numpy as np

from .vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_MEM_NONE


class KOmega:
    def __init__(self, lib, nsens, nts, navg):
        if nts % 2:
            raise ValueError("data length must be even")
        self.lib, self.nsens, self.nts, self.navg = lib, nsens, nts, navg
        self.nfreq = nts // 2 + 1
        L = lib
        self.cm_freq = L.cmcreate_f(nsens, self.nfreq, VSIP_COL, VSIP_MEM_NONE)
        self.rm_freq = L.mrealview_f(self.cm_freq)
        self.m_gram = L.mcreate_f(nsens, self.nfreq, VSIP_COL, VSIP_MEM_NONE)
        self.rcfftm = L.rcfftmop_create_f(nsens, nts, 1.0, VSIP_ROW, 0, 0)
        # kw.c:52 passes (..., VSIP_FFT_FWD, 1, VSIP_COL, ...): scale -1, dir +1
        self.ccfftm = L.ccfftmip_create_f(nsens, self.nfreq, float(VSIP_FFT_FWD), 1, VSIP_COL, 0, 0)
        self.ts_taper = L.vcreate_hanning_f(nts, 0)
        self.array_taper = L.vcreate_hanning_f(nsens, 0)
        for h in (self.cm_freq, self.rm_freq, self.m_gram, self.rcfftm, self.ccfftm, self.ts_taper,
                  self.array_taper):
            if not h:
                raise RuntimeError("k-omega initialization failed")

    def zero(self):
        """kw_zero (kw.c:3-5)."""
        self.lib.mfill_f(0.0, self.m_gram)

    def step(self, m_data):
        """komega (kw.c:7-25): m_data is an nsens x nts real view; it is
        tapered in place, as the reference does."""
        L = self.lib
        L.vmmul_f(self.ts_taper, m_data, VSIP_ROW, m_data)
        L.vmmul_f(self.array_taper, m_data, VSIP_COL, m_data)
        L.rcfftmop_f(self.rcfftm, m_data, self.cm_freq)
        L.ccfftmip_f(self.ccfftm, self.cm_freq)
        L.mcmagsq_f(self.cm_freq, self.rm_freq)
        L.smmul_f(1.0 / float(self.navg), self.rm_freq, self.rm_freq)
        L.madd_f(self.rm_freq, self.m_gram, self.m_gram)

    def step_fused(self, m_data):
        """The same step in two launches (libvsip_amd.so only); m_data is not
        modified."""
        rc = self.lib.amd_komega_f(self.rcfftm, self.ccfftm, self.ts_taper, self.array_taper, m_data,
                                   self.cm_freq, self.m_gram, 1.0 / float(self.navg))
        if rc:
            raise RuntimeError("vsip_amd_komega_f: %s" % self.lib.last_error())

    def gram(self):
        """The gram as an nsens x nfreq float32 array."""
        out = np.empty((self.nfreq, self.nsens), np.float32)      # COL order
        self.lib.mcopyto_user_f(self.m_gram, VSIP_COL, out.ctypes.data)
        return out.T.copy()

    def destroy(self):
        """kw_fin (kw.c:68-78)."""
        L = self.lib
        L.malldestroy_f(self.m_gram)
        L.fftm_destroy_f(self.rcfftm)
        L.fftm_destroy_f(self.ccfftm)
        L.valldestroy_f(self.ts_taper)
        L.valldestroy_f(self.array_taper)
        L.mdestroy_f(self.rm_freq)
        L.cmalldestroy_f(self.cm_freq)


def data_matrix(lib, a):
    """An nsens x nts ROW-major library matrix holding float32 array a."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    m = lib.mcreate_f(a.shape[0], a.shape[1], VSIP_ROW, VSIP_MEM_NONE)
    lib.mcopyfrom_user_f(a.ctypes.data, VSIP_ROW, m)
    return m


def read_matrix(lib, m, shape):
    out = np.empty(shape, np.float32)
    lib.mcopyto_user_f(m, VSIP_ROW, out.ctypes.data)
    return out


def plane_waves(nsens, nts, waves, seed=0, noise=0.0):
    """nsens x nts samples of sum_w a cos(2 pi (f t / nts + k s / nsens)) for
    waves (f, k, a) at integer bins, plus uniform noise (splitmix64)."""
    from . import datagen
    s = np.arange(nsens)[:, None]
    t = np.arange(nts)[None, :]
    x = np.zeros((nsens, nts))
    for f, k, a in waves:
        x += a * np.cos(2 * np.pi * (f * t / nts + k * s / nsens))
    if noise:
        x += noise * datagen.real_data(seed, (nsens, nts), dist="uniform", dtype=np.float64)
    return x.astype(np.float32)
