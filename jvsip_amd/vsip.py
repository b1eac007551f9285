"""ctypes binding of the VSIPL multiple-FFT C API.

`VsipLib(path)` binds the vsip_* entry points declared in include/vsip.h from
any shared library that exports them: the MI355X library built in
jvsip_amd/lib/libvsip_amd.so (see `load()`), or a build of rrjudd/jvsip's own
C sources (used only by the tests and the CPU baseline).  Handles are opaque
pointers, exactly as in C; arguments and results keep the C meaning.

`load()` returns the MI355X library and raises if it has not been built: there
is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# VSIP_AMD_LIB: another build of the library (for example to compare two builds)
LIB_PATH = os.environ.get("VSIP_AMD_LIB") or os.path.join(_HERE, "lib", "libvsip_amd.so")

c_vi = ctypes.c_ulong      # vsip_length / vsip_offset / vsip_index
c_st = ctypes.c_long       # vsip_stride
c_bl = ctypes.c_int        # vsip_scalar_bl, enums
P = ctypes.c_void_p

VSIP_ROW, VSIP_COL = 0, 1
VSIP_FFT_FWD, VSIP_FFT_INV = -1, 1
VSIP_FFT_IP, VSIP_FFT_OP = 0, 1
VSIP_MEM_NONE = 0
VSIP_ALG_TIME = 1
VSIP_SUPPORT_FULL, VSIP_SUPPORT_SAME, VSIP_SUPPORT_MIN = 0, 1, 2
VSIP_BIASED, VSIP_UNBIASED = 0, 1
VSIP_NONSYM, VSIP_SYM_EVEN_LEN_ODD, VSIP_SYM_EVEN_LEN_EVEN = 0, 1, 2


class cscalar_f(ctypes.Structure):
    _fields_ = [("r", ctypes.c_float), ("i", ctypes.c_float)]


class cscalar_d(ctypes.Structure):
    _fields_ = [("r", ctypes.c_double), ("i", ctypes.c_double)]


class scalar_mi(ctypes.Structure):
    _fields_ = [("r", c_vi), ("c", c_vi)]


class fftm_attr_f(ctypes.Structure):
    _fields_ = [("input", scalar_mi), ("output", scalar_mi), ("place", c_bl),
                ("scale", ctypes.c_float), ("dir", c_bl), ("major", c_bl)]


class fftm_attr_d(ctypes.Structure):
    _fields_ = [("input", scalar_mi), ("output", scalar_mi), ("place", c_bl),
                ("scale", ctypes.c_double), ("dir", c_bl), ("major", c_bl)]


class amd_request(ctypes.Structure):
    _fields_ = [("t_arrive", ctypes.c_double), ("rows", ctypes.c_long), ("n", ctypes.c_long),
                ("dir", ctypes.c_int), ("scale", ctypes.c_float), ("in_offset", ctypes.c_long),
                ("out_offset", ctypes.c_long)]


class amd_result(ctypes.Structure):
    _fields_ = [("t_admit", ctypes.c_double), ("t_done", ctypes.c_double), ("batch", ctypes.c_int),
                ("stream", ctypes.c_int)]


class fft_attr_f(ctypes.Structure):
    _fields_ = [("input", c_vi), ("output", c_vi), ("place", c_bl), ("scale", ctypes.c_float), ("dir", c_bl)]


class fft_attr_d(ctypes.Structure):
    _fields_ = [("input", c_vi), ("output", c_vi), ("place", c_bl), ("scale", ctypes.c_double), ("dir", c_bl)]


class corr1d_attr(ctypes.Structure):
    _fields_ = [("ref_len", c_vi), ("data_len", c_vi), ("support", c_bl), ("lag_len", c_vi)]


class conv1d_attr(ctypes.Structure):
    _fields_ = [("kernel_len", c_vi), ("symm", c_bl), ("data_len", c_vi), ("support", c_bl), ("out_len", c_vi),
                ("decimation", ctypes.c_int)]


class vattr(ctypes.Structure):
    _fields_ = [("offset", c_vi), ("stride", c_st), ("length", c_vi), ("block", P)]


class cmattr(ctypes.Structure):
    _fields_ = [("offset", c_vi), ("row_stride", c_st), ("row_length", c_vi),
                ("col_stride", c_st), ("col_length", c_vi), ("block", P)]


def _sig(t):
    """(name suffix -> (restype, argtypes)) for element type t in {f, d}."""
    sc = ctypes.c_float if t == "f" else ctypes.c_double
    ps = ctypes.POINTER(sc)
    cs = cscalar_f if t == "f" else cscalar_d
    attr = fftm_attr_f if t == "f" else fftm_attr_d
    attr1 = fft_attr_f if t == "f" else fft_attr_d
    return {
        "blockbind": (P, [P, c_vi, c_bl]),
        "blockcreate": (P, [c_vi, c_bl]),
        "blockadmit": (ctypes.c_int, [P, c_bl]),
        "blockrelease": (P, [P, c_bl]),
        "blockdestroy": (None, [P]),
        "cblockbind": (P, [P, P, c_vi, c_bl]),
        "cblockcreate": (P, [c_vi, c_bl]),
        "cblockadmit": (ctypes.c_int, [P, c_bl]),
        "cblockrelease": (None, [P, c_bl, ctypes.POINTER(P), ctypes.POINTER(P)]),
        "cblockdestroy": (None, [P]),
        "cblockfind": (None, [P, ctypes.POINTER(P), ctypes.POINTER(P)]),
        "cblockrebind": (None, [P, P, P, ctypes.POINTER(P), ctypes.POINTER(P)]),
        "blockfind": (P, [P]),
        "blockrebind": (P, [P, P]),
        "cmbind": (P, [P, c_vi, c_st, c_vi, c_st, c_vi]),
        "cmcreate": (P, [c_vi, c_vi, c_bl, c_bl]),
        "cmsubview": (P, [P, c_vi, c_vi, c_vi, c_vi]),
        "cmdestroy": (P, [P]),
        "cmalldestroy": (None, [P]),
        "cmget": (cs, [P, c_vi, c_vi]),
        "cmput": (None, [P, c_vi, c_vi, cs]),
        "cmcopyfrom_user": (None, [P, P, c_bl, P]),
        "cmcopyto_user": (None, [P, c_bl, P, P]),
        "cmgetattrib": (None, [P, ctypes.POINTER(cmattr)]),
        "cmtransview": (P, [P]),
        "mbind": (P, [P, c_vi, c_st, c_vi, c_st, c_vi]),
        "mcreate": (P, [c_vi, c_vi, c_bl, c_bl]),
        "mdestroy": (P, [P]),
        "malldestroy": (None, [P]),
        "mcopyfrom_user": (None, [P, c_bl, P]),
        "mcopyto_user": (None, [P, c_bl, P]),
        "msubview": (P, [P, c_vi, c_vi, c_vi, c_vi]),
        "mget": (sc, [P, c_vi, c_vi]),
        "mput": (None, [P, c_vi, c_vi, sc]),
        "ccfftmip_create": (P, [c_vi, c_vi, sc, c_bl, c_bl, ctypes.c_uint, c_bl]),
        "ccfftmop_create": (P, [c_vi, c_vi, sc, c_bl, c_bl, ctypes.c_uint, c_bl]),
        "rcfftmop_create": (P, [c_vi, c_vi, sc, c_bl, ctypes.c_uint, c_bl]),
        "crfftmop_create": (P, [c_vi, c_vi, sc, c_bl, ctypes.c_uint, c_bl]),
        "ccfftmip": (None, [P, P]),
        "ccfftmop": (None, [P, P, P]),
        "rcfftmop": (None, [P, P, P]),
        "crfftmop": (None, [P, P, P]),
        "fftm_destroy": (ctypes.c_int, [P]),
        "fftm_getattr": (None, [P, ctypes.POINTER(attr)]),
        "cmsub": (None, [P, P, P]),
        "cmmeansqval": (sc, [P]),
        "cvsub": (None, [P, P, P]),
        "cvmeansqval": (sc, [P]),
        "rscvmul": (None, [sc, P, P]),
        "rscmmul": (None, [sc, P, P]),
        # real element operations and part views (k-omega beamformer)
        "vmmul": (None, [P, P, c_bl, P]),
        "mcmagsq": (None, [P, P]),
        "smmul": (None, [sc, P, P]),
        "madd": (None, [P, P, P]),
        "mfill": (None, [sc, P]),
        "vcreate_hanning": (P, [c_vi, c_bl]),
        "mrealview": (P, [P]),
        "mimagview": (P, [P]),
        # random numbers
        "randu": (sc, [P]),
        "randn": (sc, [P]),
        "crandu": (cs, [P]),
        "crandn": (cs, [P]),
        "vrandu": (None, [P, P]),
        "vrandn": (None, [P, P]),
        "cvrandu": (None, [P, P]),
        "cvrandn": (None, [P, P]),
        "mrandu": (None, [P, P]),
        "mrandn": (None, [P, P]),
        "cmrandu": (None, [P, P]),
        "cmrandn": (None, [P, P]),
        # vector views and the 1-D FFT
        "vbind": (P, [P, c_vi, c_st, c_vi]),
        "vcreate": (P, [c_vi, c_bl]),
        "vdestroy": (P, [P]),
        "valldestroy": (None, [P]),
        "vget": (sc, [P, c_vi]),
        "vput": (None, [P, c_vi, sc]),
        "vgetattrib": (None, [P, ctypes.POINTER(vattr)]),
        "vcopyfrom_user": (None, [P, P]),
        "vcopyto_user": (None, [P, P]),
        "cvbind": (P, [P, c_vi, c_st, c_vi]),
        "cvcreate": (P, [c_vi, c_bl]),
        "cvdestroy": (P, [P]),
        "cvalldestroy": (None, [P]),
        "cvget": (cs, [P, c_vi]),
        "cvput": (None, [P, c_vi, cs]),
        "cvgetattrib": (None, [P, ctypes.POINTER(vattr)]),
        "cvcopyfrom_user": (None, [P, P, P]),
        "cvcopyto_user": (None, [P, P, P]),
        "ccfftip_create": (P, [c_vi, sc, c_bl, ctypes.c_uint, c_bl]),
        "ccfftop_create": (P, [c_vi, sc, c_bl, ctypes.c_uint, c_bl]),
        "rcfftop_create": (P, [c_vi, sc, ctypes.c_uint, c_bl]),
        "crfftop_create": (P, [c_vi, sc, ctypes.c_uint, c_bl]),
        "ccfftip": (None, [P, P]),
        "ccfftop": (None, [P, P, P]),
        "rcfftop": (None, [P, P, P]),
        "crfftop": (None, [P, P, P]),
        "fft_destroy": (ctypes.c_int, [P]),
        "fft_getattr": (None, [P, ctypes.POINTER(attr1)]),
        # 1-D correlation
        "corr1d_create": (P, [c_vi, c_vi, c_bl, ctypes.c_uint, c_bl]),
        "ccorr1d_create": (P, [c_vi, c_vi, c_bl, ctypes.c_uint, c_bl]),
        "corr1d_destroy": (ctypes.c_int, [P]),
        "ccorr1d_destroy": (ctypes.c_int, [P]),
        "corr1d_getattr": (None, [P, ctypes.POINTER(corr1d_attr)]),
        "ccorr1d_getattr": (None, [P, ctypes.POINTER(corr1d_attr)]),
        "correlate1d": (None, [P, c_bl, P, P, P]),
        "ccorrelate1d": (None, [P, c_bl, P, P, P]),
        # 1-D convolution
        "conv1d_create": (P, [P, c_bl, c_vi, ctypes.c_int, c_bl, ctypes.c_uint, c_bl]),
        "conv1d_destroy": (ctypes.c_int, [P]),
        "conv1d_getattr": (None, [P, ctypes.POINTER(conv1d_attr)]),
        "convolve1d": (None, [P, P, P]),
    }


class VsipLib:
    """The vsip_* API of one shared library; attribute `x_f` is vsip_x_f."""

    def __init__(self, path, required=True):
        if not os.path.exists(path):
            raise FileNotFoundError(
                "%s is not built (run `python -c \"import __graft_entry__ as g; g.build()\"`)" % path)
        self.path = path
        self.dll = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        self.missing = []
        for t in ("f", "d"):
            for name, (res, args) in _sig(t).items():
                full = "vsip_%s_%s" % (name, t)
                try:
                    fn = getattr(self.dll, full)
                except AttributeError:
                    self.missing.append(full)
                    continue
                fn.restype = res
                fn.argtypes = args
                setattr(self, "%s_%s" % (name, t), fn)
        for name in ("vsip_init", "vsip_finalize"):
            fn = getattr(self.dll, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [P]
            setattr(self, name[5:], fn)
        for name, res, args in (("vsip_randcreate", P, [c_vi, c_vi, c_vi, c_bl]),
                                ("vsip_randdestroy", ctypes.c_int, [P])):
            fn = getattr(self.dll, name, None)
            if fn is None:
                self.missing.append(name)
                continue
            fn.restype = res
            fn.argtypes = args
            setattr(self, name[5:], fn)
        self.init(None)


class AmdLib(VsipLib):
    """libvsip_amd.so: the VSIPL API plus the vsip_amd_* extensions."""

    def __init__(self, path=LIB_PATH):
        super().__init__(path)
        d = self.dll
        ext = {
            "vsip_amd_stream": (P, []),
            "vsip_amd_sync": (ctypes.c_int, []),
            "vsip_amd_last_error": (ctypes.c_char_p, []),
            "vsip_amd_clear_error": (None, []),
            "vsip_amd_bytes_in_use": (ctypes.c_size_t, []),
            "vsip_amd_set_hbm_budget": (None, [ctypes.c_size_t]),
            "vsip_amd_cblock_device_f": (P, [P, ctypes.POINTER(P)]),
            "vsip_amd_cblock_device_d": (P, [P, ctypes.POINTER(P)]),
            "vsip_amd_block_device_f": (P, [P]),
            "vsip_amd_block_device_d": (P, [P]),
            "vsip_amd_schedule": (ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                                 ctypes.c_char_p, ctypes.c_size_t]),
            "vsip_amd_fftm_describe_f": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_size_t]),
            "vsip_amd_fftm_describe_d": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_size_t]),
            "vsip_amd_batch_create_f": (P, [c_vi, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P)]),
            "vsip_amd_batch_run_f": (ctypes.c_int, [P]),
            "vsip_amd_batch_run_noalloc_f": (ctypes.c_int, [c_vi, ctypes.POINTER(P), ctypes.POINTER(P),
                                                            ctypes.POINTER(P)]),
            "vsip_amd_batch_destroy_f": (None, [P]),
            "vsip_amd_batch_info_f": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_size_t]),
            "vsip_amd_hbm_probe": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_double)]),
            "vsip_amd_col_probe": (ctypes.c_int, [ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                                  ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
            "vsip_amd_tune": (ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_char_p, ctypes.c_size_t]),
            "vsip_amd_serve_f": (ctypes.c_int, [ctypes.POINTER(amd_request), ctypes.c_long, P, P, ctypes.c_size_t,
                                                ctypes.c_int, ctypes.c_int, ctypes.POINTER(amd_result),
                                                ctypes.c_char_p, ctypes.c_size_t]),
            "vsip_amd_serve_plain_f": (ctypes.c_int, [ctypes.POINTER(amd_request), ctypes.c_long, P, P,
                                                      ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(amd_result),
                                                      ctypes.c_char_p, ctypes.c_size_t]),
            "vsip_amd_komega_f": (ctypes.c_int, [P, P, P, P, P, P, P, ctypes.c_float]),
        }
        for name, (res, args) in ext.items():
            fn = getattr(d, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name[5:], fn)

    def last_error(self):
        return self.amd_last_error().decode()

    def schedule(self, prec, n, nvec, lds_max=163840):
        """Planner output for nvec transforms of length n (no device needed)."""
        import json
        buf = ctypes.create_string_buffer(2048)
        self.amd_schedule(prec, n, nvec, lds_max, buf, 2048)
        return json.loads(buf.value.decode())

    def batch(self, plans, ins, outs=None):
        """vsip_amd_batch_create_f over Python lists of handles."""
        n = len(plans)
        arr = lambda xs: (P * n)(*xs)
        return self.amd_batch_create_f(n, arr(plans), arr(ins), arr(outs) if outs else None)

    def batch_noalloc(self, plans, ins, outs=None):
        """vsip_amd_batch_run_noalloc_f over Python lists of handles."""
        n = len(plans)
        arr = lambda xs: (P * n)(*xs)
        return self.amd_batch_run_noalloc_f(n, arr(plans), arr(ins), arr(outs) if outs else None)

    def batch_info(self, b):
        import json
        buf = ctypes.create_string_buffer(1024)
        self.amd_batch_info_f(b, buf, 1024)
        return json.loads(buf.value.decode())

    def tune(self, n, prec=4):
        import json
        buf = ctypes.create_string_buffer(1024)
        self.amd_tune(prec, n, buf, 1024)
        return json.loads(buf.value.decode())

    def serve(self, reqs, in_pool, out_pool=None, budget=64 << 30, lanes=4, max_jobs=64, plain=False,
              sync_each=True):
        """Run a request stream (list of dicts or tuples in amd_request field
        order) through the server (or the plain baseline: plain calls with a
        synchronisation after each, or back to back with sync_each=False);
        returns (results as a numpy structured array, stats dict)."""
        import json
        import numpy as np
        n = len(reqs)
        arr = (amd_request * n)()
        for i, r in enumerate(reqs):
            arr[i] = amd_request(*r) if not isinstance(r, dict) else amd_request(**r)
        res = (amd_result * n)()
        buf = ctypes.create_string_buffer(2048)
        if plain:
            rc = self.amd_serve_plain_f(arr, n, in_pool, out_pool, budget, 1 if sync_each else 0, res, buf, 2048)
        else:
            rc = self.amd_serve_f(arr, n, in_pool, out_pool, budget, lanes, max_jobs, res, buf, 2048)
        if rc:
            raise RuntimeError("serve failed (%d): %s" % (rc, self.last_error()))
        out = np.array([(r.t_admit, r.t_done, r.batch, r.stream) for r in res],
                       dtype=[("t_admit", "f8"), ("t_done", "f8"), ("batch", "i4"), ("stream", "i4")])
        return out, json.loads(buf.value.decode() or "{}")

    def describe(self, plan, t="f"):
        buf = ctypes.create_string_buffer(4096)
        getattr(self, "amd_fftm_describe_" + t)(plan, buf, 4096)
        return buf.value.decode()


_LIB = None


def load():
    """The MI355X library (built in-tree); raises if it is missing."""
    global _LIB
    if _LIB is None:
        _LIB = AmdLib(LIB_PATH)
    return _LIB
