"""Benchmark: multiple FFT throughput on one MI355X.

Workload (one step): vsip_ccfftmip_f forward (scale 1) on one 4096 x 4096
complex FP32 matrix, ROW, then vsip_ccfftmip_f inverse (scale 1/4096) on
another, inputs iid uniform[-1, 1] (seed 2).  Four device-resident matrices
(512 MiB, twice the 256 MB Infinity Cache) are visited in rotation so every
transform reads cold HBM; each matrix alternates forward and inverse visits,
so the data stay bounded.

value = algorithmic bytes / time: one read plus one write of the logical
data per transform, 268,435,456 B, i.e. 536,870,912 B per step.

Also reported:
  roofline      the kernel's own rate: HIP events on the library stream
                bracket the timed region, which holds only the 2 * steps kernel
                launches, so their average duration is (event time) / launches;
                against 8 TB/s; `traffic` is the HBM
                bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
                (profiles/traffic_ccfftmip_4096x4096.json) or null
  cpu_baseline  rrjudd/jvsip's own C (oracle/_ref/libjvsip_ref.so) on the host
                cores of this machine, on a bounded sample of the same workload
  accuracy      relative L2 of 8 output rows against jvsip's output fixture
                (tests/golden) and against numpy's double-precision FFT;
                the oracle under oracle/ is used only by the cpu_baseline leg
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
M = N = 4096
BYTES_PER_CALL = 2 * M * N * 8          # read + write of 4096 x 4096 complex64
NBUF = 4


class Hip:
    """The few HIP runtime calls needed for event timing on a given stream."""

    def __init__(self):
        self.d = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        self.d.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.d.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.d.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.d.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.d.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.d.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.d.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        self.d.hipEventSynchronize(b)
        t = ctypes.c_float()
        assert self.d.hipEventElapsedTime(ctypes.byref(t), a, b) == 0
        return t.value


def cpu_share():
    """The host cores this process may use: the CPUs of its affinity mask,
    capped by the cgroup's CPU quota (cpu.max) when one is set, and at most 16
    (one GPU's share of the host on a multi-GPU node); with the host's
    logical CPU count and CPU model, as lscpu reports them."""
    n_all = os.cpu_count() or 1
    try:
        n_aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n_aff = n_all
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    cores = n_aff if quota is None else max(1, min(n_aff, int(quota)))
    cores = min(cores, 16)
    model = "unknown"
    try:
        import subprocess
        for line in subprocess.check_output(["lscpu"], text=True).splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"cores": cores, "nproc": n_all, "affinity_cpus": n_aff, "cgroup_cpu_quota": quota,
            "lscpu_model": model,
            "policy": "threads = the affinity CPUs, capped by the cgroup quota and by 16, "
                      "one GPU's share of the host"}


def cpu_baseline(threads, steps=5, rows_single=256):
    """jvsip's C on the host: fwd+inv of the 4096 x 4096 matrix split into
    `threads` row slabs (one plan per thread; plans are independent), timed
    over `steps` steps after one warm-up; plus one core on `rows_single` rows."""
    from concurrent.futures import ThreadPoolExecutor
    from jvsip_amd import datagen
    from oracle import pyoracle
    if not pyoracle.ref_available():
        return None
    ref = pyoracle.ref()
    x = datagen.complex_data(2, (M, N), dist="uniform")
    slabs = np.array_split(np.arange(M), threads)

    def make(rows):
        buf = np.ascontiguousarray(x[rows[0]:rows[-1] + 1]).ravel().copy()
        m = len(rows)
        blk = ref.cblockbind_f(buf.ctypes.data, None, buf.size, 0)
        ref.cblockadmit_f(blk, 1)
        view = ref.cmbind_f(blk, 0, N, m, 1, N)
        fwd = ref.ccfftmip_create_f(m, N, 1.0, -1, 0, 0, 1)
        inv = ref.ccfftmip_create_f(m, N, 1.0 / N, 1, 0, 0, 1)
        return buf, blk, view, fwd, inv

    work = [make(r) for r in slabs]

    def one(w):
        ref.ccfftmip_f(w[3], w[2])
        ref.ccfftmip_f(w[4], w[2])

    with ThreadPoolExecutor(threads) as pool:
        list(pool.map(one, work))                       # warm-up
        t0 = time.perf_counter()
        for _ in range(steps):
            list(pool.map(one, work))
        dt = (time.perf_counter() - t0) / steps
    for buf, blk, view, fwd, inv in work:
        ref.fftm_destroy_f(fwd)
        ref.fftm_destroy_f(inv)
        ref.cmdestroy_f(view)
        ref.cblockdestroy_f(blk)
    # one core, a row sample
    w = make(np.arange(rows_single))
    one(w)
    t0 = time.perf_counter()
    one(w)
    dt1 = time.perf_counter() - t0
    gbs_all = 2 * BYTES_PER_CALL / dt / 1e9
    gbs_one = 2 * (BYTES_PER_CALL * rows_single // M) / dt1 / 1e9
    try:
        cpu = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        cpu = "unknown"
    per_config = cpu_configs(ref)
    return {
        "value": round(gbs_all, 4), "unit": "GB/s", "cores": threads, "kind": "reference",
        "sample": "jvsip vsip_ccfftmip_f fwd+inv on the full 4096x4096 matrix as %d row slabs "
                  "(one plan per thread), %d timed steps after 1 warm-up; %.3f s per step" % (threads, steps, dt),
        "single_core": {"value": round(gbs_one, 4), "unit": "GB/s", "cores": 1,
                        "sample": "fwd+inv on %d rows x 4096, %.3f s" % (rows_single, dt1)},
        "cpu_model": cpu,
        "configs": per_config,
    }


def cpu_configs(ref):
    """jvsip's C, one core, one call each (after one untimed call) on the
    other configurations' shapes: seconds per call, to set beside the GPU
    figures of the same shapes."""
    from jvsip_amd import datagen

    def timed(fn):
        fn()
        t0 = time.perf_counter()
        fn()
        return round(time.perf_counter() - t0, 5)

    out = {}
    for name, m, n, major, parent in (("c3_col_1920x1000", 1920, 1000, 1, 2000),
                                      ("c3_row_1024x2187", 1024, 2187, 0, 4374),
                                      ("c3_row_512x1001", 512, 1001, 0, 2002)):
        x = datagen.complex_data(3, (m * parent,))
        blk = ref.cblockbind_f(x.ctypes.data, None, m * parent, 0)
        ref.cblockadmit_f(blk, 1)
        xv = ref.cmbind_f(blk, 0, parent, m, 1, n)
        yb = np.zeros(m * n, np.complex64)
        yblk = ref.cblockbind_f(yb.ctypes.data, None, m * n, 0)
        ref.cblockadmit_f(yblk, 1)
        yv = ref.cmbind_f(yblk, 0, n, m, 1, n)
        p = ref.ccfftmop_create_f(m, n, 1.0, -1, major, 0, 1)
        out[name] = {"s_per_call": timed(lambda: ref.ccfftmop_f(p, xv, yv))}
        ref.fftm_destroy_f(p)
        ref.cmdestroy_f(xv)
        ref.cmdestroy_f(yv)
        ref.cblockdestroy_f(blk)
        ref.cblockdestroy_f(yblk)
    xd = datagen.complex_data(3, (2048 * 2048,), dtype=np.complex128)
    blk = ref.cblockbind_d(xd.ctypes.data, None, xd.size, 0)
    ref.cblockadmit_d(blk, 1)
    v = ref.cmbind_d(blk, 0, 2048, 2048, 1, 2048)
    p = ref.ccfftmip_create_d(2048, 2048, 1.0, -1, 0, 0, 1)
    out["c3_fp64_2048x2048"] = {"s_per_call": timed(lambda: ref.ccfftmip_d(p, v))}
    ref.fftm_destroy_d(p)
    ref.cmdestroy_d(v)
    ref.cblockdestroy_d(blk)
    xr = datagen.real_data(2, (8192 * 2048,))
    rb = ref.blockbind_f(xr.ctypes.data, xr.size, 0)
    ref.blockadmit_f(rb, 1)
    xv = ref.mbind_f(rb, 0, 2048, 8192, 1, 2048)
    z = np.zeros(8192 * 1025, np.complex64)
    zb = ref.cblockbind_f(z.ctypes.data, None, z.size, 0)
    ref.cblockadmit_f(zb, 1)
    zv = ref.cmbind_f(zb, 0, 1025, 8192, 1, 1025)
    p = ref.rcfftmop_create_f(8192, 2048, 1.0, 0, 0, 1)
    out["c2_rcfftmop_8192x2048"] = {"s_per_call": timed(lambda: ref.rcfftmop_f(p, xv, zv))}
    ref.fftm_destroy_f(p)
    ref.mdestroy_f(xv)
    ref.cmdestroy_f(zv)
    ref.blockdestroy_f(rb)
    ref.cblockdestroy_f(zb)
    from jvsip_amd import komega as kw
    k = kw.KOmega(ref, 1024, 8192, 8)
    k.zero()
    d = kw.data_matrix(ref, kw.plane_waves(1024, 8192, [(1024, 3, 1.0)], seed=7, noise=0.5))
    out["f3_komega_1024x8192"] = {"s_per_call": timed(lambda: k.step(d))}
    ref.malldestroy_f(d)
    k.destroy()
    for cplx in (True, False):
        c = _corr_setup(ref, cplx, 1024, 1 << 20)
        out[("ccorrelate1d" if cplx else "correlate1d") + "_f_1024x2^20"] = {"s_per_call": timed(c["call"])}
        c["cleanup"]()
    c = _conv_setup(ref, 1024, 1 << 20)
    out["convolve1d_f_1024x2^20"] = {"s_per_call": timed(c["call"])}
    c["cleanup"]()
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def fft64(x, axis=-1):
    """numpy's FFT in double precision (the accuracy reference of the bench)"""
    return np.fft.fft(np.asarray(x, dtype=np.complex128), axis=axis)


def accuracy(lib, view_fn):
    """Forward transform of the seed-2 input: 8 rows against jvsip's output
    (tests/golden fixture) and numpy's FP64 FFT."""
    from jvsip_amd import datagen, fftm
    f = np.load(os.path.join(ROOT, "tests", "golden", "jvsip_c2_4096x4096_fwd.npz"))
    x = datagen.complex_data(2, (M, N), dist="uniform")
    y = fftm.ccfftm(lib, x.ravel(), fftm.dense_view(M, N)).reshape(M, N)
    rows = f["rows"]
    return {"rel_l2_vs_jvsip": rel_l2(y[rows], f["out_rows"]),
            "rel_l2_vs_fp64": rel_l2(y[rows], fft64(x[rows], axis=1)),
            "rows": [int(r) for r in rows]}



# ---------------------------------------------------------------------------
# the other configurations of the workload set, each timed on the device and
# checked for accuracy
# ---------------------------------------------------------------------------

def _dev_view(lib, x, major=0):
    """device matrix (vsip_cmcreate_f) holding host array x (ROW-ordered data)"""
    m, n = x.shape
    v = lib.cmcreate_f(m, n, major, 0)
    assert v, lib.last_error()
    xx = np.ascontiguousarray(x, dtype=np.complex64)
    lib.cmcopyfrom_user_f(xx.ctypes.data, None, 0, v)
    return v


def _time_calls(lib, hip, fn, reps, warm_s=0.03):
    """wall and stream-event time per call of fn() over reps back-to-back
    calls, after at least warm_s seconds of the same calls (the device reaches
    its sustained clocks only under load)"""
    stream = lib.amd_stream()
    t0 = time.perf_counter()
    fn()
    while time.perf_counter() - t0 < warm_s:
        fn()
        lib.amd_sync()
    lib.amd_sync()
    e0, e1 = hip.event(), hip.event()
    t0 = time.perf_counter()
    hip.record(e0, stream)
    for _ in range(reps):
        fn()
    hip.record(e1, stream)
    lib.amd_sync()
    wall = (time.perf_counter() - t0) / reps
    return wall, hip.elapsed_ms(e0, e1) * 1e-3 / reps


_CEIL = {}


def size_ceiling(lib, bytes_moved):
    """The in-place read+write probe's rate at the same bytes per launch
    (4 buffers in rotation, back-to-back launches timed with stream events):
    what any single launch moving this many bytes reaches on this device,
    launch overhead included.  Small transforms are bounded by it rather than
    by the 8 TB/s peak."""
    half = max(32768, (int(bytes_moved) // 2) // 32768 * 32768)
    if half not in _CEIL:
        rmw, rd = ctypes.c_double(), ctypes.c_double()
        lib.amd_hbm_probe(half, 40, ctypes.byref(rmw), ctypes.byref(rd))
        lib.amd_hbm_probe(half, 40, ctypes.byref(rmw), ctypes.byref(rd))
        _CEIL[half] = rmw.value
    return round(_CEIL[half], 1)


def config_c1(lib, hip):
    """plain vsip_ccfftmip_f forward, ROW, 64 x 1024, scale 1, N(0,1) seed 1"""
    from jvsip_amd import datagen, fftm
    x = datagen.complex_data(1, (64, 1024))
    v = _dev_view(lib, x)
    p = lib.ccfftmip_create_f(64, 1024, 1.0, -1, 0, 0, 1)
    wall, gpu = _time_calls(lib, hip, lambda: lib.ccfftmip_f(p, v), 200)
    # one call from a fresh copy, with synchronisation, for the latency figure
    lib.cmcopyfrom_user_f(np.ascontiguousarray(x).ctypes.data, None, 0, v)
    t0 = time.perf_counter()
    lib.ccfftmip_f(p, v)
    lib.amd_sync()
    lat = time.perf_counter() - t0
    y = np.empty((64, 1024), np.complex64)
    lib.cmcopyto_user_f(v, 0, y.ctypes.data, None)
    f = np.load(os.path.join(ROOT, "tests", "golden", "jvsip_c1_64x1024.npz"))
    out = {"shape": "64x1024", "us_per_call_back_to_back_python": round(wall * 1e6, 2),
           "us_per_call_kernel_stream": round(gpu * 1e6, 2), "us_call_plus_sync": round(lat * 1e6, 2),
           "gb_per_s": round(2 * 64 * 1024 * 8 / gpu / 1e9, 1),
           "size_matched_ceiling_gb_per_s": size_ceiling(lib, 2 * 64 * 1024 * 8),
           "rel_l2_vs_jvsip": rel_l2(y, f["out"]),
           "rel_l2_vs_fp64": rel_l2(y, fft64(x, axis=1))}
    lib.fftm_destroy_f(p)
    lib.cmalldestroy_f(v)
    out["c_host_per_call"] = percall_c()
    return out


def percall_c():
    """The host cost of a plain call from C (jvsip_amd/lib/bench_percall, a
    child process): vsip_ccfftmip_f 64 x 1024 back to back against launches
    of an empty kernel with a 464-byte argument on the same stream."""
    import subprocess
    exe = os.path.join(ROOT, "jvsip_amd", "lib", "bench_percall")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe, "4000"], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return {"error": "exit %d: %s" % (r.returncode, r.stderr[-200:])}
    d = json.loads(r.stdout.strip().splitlines()[-1])
    d["call_minus_empty_launch_us"] = round(d["ccfftmip_f_64x1024_host_us"] - d["empty_kernel_host_us"], 3)
    return d


def config_host_buffers(lib, reps=3):
    """The drop-in's host-buffer form of the headline transform: a 4096 x 4096
    user array (pageable numpy memory) bound with vsip_cblockbind_f, then
    vsip_cblockadmit_f(update) (upload), vsip_ccfftmip_f forward,
    vsip_cblockrelease_f(update) (download).  Reported beside the headline,
    which keeps its inputs resident in HBM."""
    from jvsip_amd import datagen
    x = datagen.complex_data(2, (M * N,), dist="uniform")
    buf = x.copy()
    blk = lib.cblockbind_f(buf.ctypes.data, None, M * N, 0)
    v = lib.cmbind_f(blk, 0, N, M, 1, N)
    p = lib.ccfftmip_create_f(M, N, 1.0, -1, 0, 0, 1)
    assert blk and v and p, lib.last_error()
    r, i = ctypes.c_void_p(), ctypes.c_void_p()
    t_admit = t_fft = t_rel = 0.0
    for k in range(reps + 1):
        t0 = time.perf_counter()
        assert lib.cblockadmit_f(blk, 1) == 0, lib.last_error()
        t1 = time.perf_counter()
        lib.ccfftmip_f(p, v)
        lib.amd_sync()
        t2 = time.perf_counter()
        lib.cblockrelease_f(blk, 1, ctypes.byref(r), ctypes.byref(i))
        t3 = time.perf_counter()
        if k:
            t_admit += t1 - t0
            t_fft += t2 - t1
            t_rel += t3 - t2
    total = (t_admit + t_fft + t_rel) / reps
    out = {"shape": "4096x4096 complex FP32, pageable user array",
           "admit_ms": round(t_admit / reps * 1e3, 2), "fft_ms": round(t_fft / reps * 1e3, 3),
           "release_ms": round(t_rel / reps * 1e3, 2), "total_ms": round(total * 1e3, 2),
           "gb_per_s_pcie_inclusive": round(BYTES_PER_CALL / total / 1e9, 2)}
    lib.fftm_destroy_f(p)
    lib.cmdestroy_f(v)
    lib.cblockdestroy_f(blk)
    return out


def config_c2_real(lib, hip):
    """vsip_rcfftmop_f 8192 x 2048 real N(0,1) then vsip_crfftmop_f (scale 1/2048) back"""
    from jvsip_amd import datagen
    M, N = 8192, 2048
    x = datagen.real_data(2, (M, N))
    xv = lib.mcreate_f(M, N, 0, 0)
    lib.mcopyfrom_user_f(np.ascontiguousarray(x).ctypes.data, 0, xv)
    zv = lib.cmcreate_f(M, N // 2 + 1, 0, 0)
    bv = lib.mcreate_f(M, N, 0, 0)
    rc = lib.rcfftmop_create_f(M, N, 1.0, 0, 0, 1)
    cr = lib.crfftmop_create_f(M, N, 1.0 / N, 0, 0, 1)
    _, t_rc = _time_calls(lib, hip, lambda: lib.rcfftmop_f(rc, xv, zv), 20)
    _, t_cr = _time_calls(lib, hip, lambda: lib.crfftmop_f(cr, zv, bv), 20)
    lib.rcfftmop_f(rc, xv, zv)
    lib.crfftmop_f(cr, zv, bv)
    back = np.empty((M, N), np.float32)
    lib.mcopyto_user_f(bv, 0, back.ctypes.data)
    z = np.empty((M, N // 2 + 1), np.complex64)
    lib.cmcopyto_user_f(zv, 0, z.ctypes.data, None)
    rows = np.arange(0, M, 1024)
    zz = np.fft.rfft(x[rows].astype(np.float64), axis=1)
    bytes_ = M * N * 4 + M * (N // 2 + 1) * 8
    out = {"shape": "8192x2048 real", "rc_us": round(t_rc * 1e6, 2), "cr_us": round(t_cr * 1e6, 2),
           "rc_gb_per_s": round(bytes_ / t_rc / 1e9, 1), "cr_gb_per_s": round(bytes_ / t_cr / 1e9, 1),
           "rc_rel_l2_vs_fp64": rel_l2(z[rows], zz), "round_trip_rel_l2": rel_l2(back, x)}
    for h in (rc, cr):
        lib.fftm_destroy_f(h)
    for v in (xv, bv):
        lib.malldestroy_f(v)
    lib.cmalldestroy_f(zv)
    return out


def config_p2_rows(lib, hip):
    """large batches of power-of-two rows (the pipelined kernels up to 4096
    points, one-row kernels beyond; set VSIP_AMD_P2_PIPE=0 for the
    one-group-per-workgroup ones throughout): c2c forward +
    inverse (scale 1/N) alternating over 4 matrices in rotation, and
    rcfftmop / crfftmop (scale 1/N) per call; microseconds per launch and
    algorithmic GB/s.  The first 64 rows hold U[-1,1] data, the rest of each
    matrix is zero."""
    from jvsip_amd import datagen
    out = {"pipelined": os.environ.get("VSIP_AMD_P2_PIPE", "1") != "0"}
    for m, n in ((65536, 256), (32768, 512), (16384, 1024), (8192, 2048), (65536, 1024), (2048, 8192),
                 (1024, 16384)):
        x = datagen.complex_data(2, (64, n), dist="uniform").ravel()
        views = []
        for _ in range(NBUF):
            v = lib.cmcreate_f(m, n, 0, 0)
            assert v, lib.last_error()
            sub = lib.cmsubview_f(v, 0, 0, 64, n)
            lib.cmcopyfrom_user_f(x.ctypes.data, None, 0, sub)
            lib.cmdestroy_f(sub)
            views.append(v)
        fwd = lib.ccfftmip_create_f(m, n, 1.0, -1, 0, 0, 1)
        inv = lib.ccfftmip_create_f(m, n, 1.0 / n, 1, 0, 0, 1)
        k = [0]

        def pair():
            lib.ccfftmip_f(fwd, views[k[0] % NBUF])
            lib.ccfftmip_f(inv, views[(k[0] + NBUF // 2) % NBUF])
            k[0] += 1
        _, t = _time_calls(lib, hip, pair, 40)
        out["c2c_%dx%d" % (m, n)] = {"us": round(t / 2 * 1e6, 2), "gb_per_s": round(16 * m * n / (t / 2) / 1e9, 1)}
        lib.fftm_destroy_f(fwd)
        lib.fftm_destroy_f(inv)
        for v in views:
            lib.cmalldestroy_f(v)
    for m, n in ((32768, 512), (16384, 1024), (8192, 2048), (4096, 4096), (2048, 8192)):
        xv = lib.mcreate_f(m, n, 0, 0)
        zv = lib.cmcreate_f(m, n // 2 + 1, 0, 0)
        sub = lib.msubview_f(xv, 0, 0, 64, n)
        xr = datagen.real_data(2, (64, n))
        lib.mcopyfrom_user_f(np.ascontiguousarray(xr).ctypes.data, 0, sub)
        lib.mdestroy_f(sub)
        rc = lib.rcfftmop_create_f(m, n, 1.0, 0, 0, 1)
        cr = lib.crfftmop_create_f(m, n, 1.0 / n, 0, 0, 1)
        _, t_rc = _time_calls(lib, hip, lambda: lib.rcfftmop_f(rc, xv, zv), 40)
        _, t_cr = _time_calls(lib, hip, lambda: lib.crfftmop_f(cr, zv, xv), 40)
        b = m * n * 4 + m * (n // 2 + 1) * 8
        out["real_%dx%d" % (m, n)] = {"rc_us": round(t_rc * 1e6, 2), "rc_gb_per_s": round(b / t_rc / 1e9, 1),
                                      "cr_us": round(t_cr * 1e6, 2), "cr_gb_per_s": round(b / t_cr / 1e9, 1)}
        lib.fftm_destroy_f(rc)
        lib.fftm_destroy_f(cr)
        lib.malldestroy_f(xv)
        lib.cmalldestroy_f(zv)
    assert not lib.last_error(), lib.last_error()
    return out


def col_pattern(lib, plan, major, m, n, parent, t):
    """For a COL plan: the rate of its access pattern with no transform in
    it (vsip_amd_col_probe: the plan's columns per workgroup, rows `parent`
    apart on the load side and n apart on the store side) and the plan's
    fraction of it; nothing for ROW plans."""
    passes = json.loads(lib.describe(plan))["passes"]
    if major != 1 or len(passes) != 1:
        return {}
    vpb = int(passes[0]["across"][0])
    g = ctypes.c_double(0.0)
    if lib.amd_col_probe(m, n, parent, n, vpb, 40, ctypes.byref(g)) or g.value <= 0:
        return {}
    lib.amd_col_probe(m, n, parent, n, vpb, 40, ctypes.byref(g))
    return {"pattern_ceiling_gb_per_s": round(g.value, 1), "pattern_columns_per_workgroup": vpb,
            "frac_of_pattern_ceiling": round(2 * m * n * 8 / t / 1e9 / g.value, 3)}


def config_c3(lib, hip):
    """strided / mixed-radix ccfftmop_f (parent rows 2x the row length) and the FP64 2048^2 case"""
    from jvsip_amd import datagen, fftm
    out = {}
    cases = [("col_1920x1000", "jvsip_c3_col_1920x1000.npz", 1920, 1000, 1, 2000),
             ("row_1024x2187", "jvsip_c3_row_1024x2187.npz", 1024, 2187, 0, 4374),
             ("row_512x1001", "jvsip_c3_row_512x1001.npz", 512, 1001, 0, 2002)]
    for name, fx, m, n, major, parent in cases:
        f = np.load(os.path.join(ROOT, "tests", "golden", fx))
        x = datagen.complex_data(3, (m * parent,))
        blk = lib.cblockcreate_f(m * parent, 0)
        pv = lib.cmbind_f(blk, 0, parent, m, 1, parent)
        lib.cmcopyfrom_user_f(x.ctypes.data, None, 0, pv)
        xv = lib.cmsubview_f(pv, 0, 0, m, n)              # left half of the parent
        yv = lib.cmcreate_f(m, n, 0, 0)
        p = lib.ccfftmop_create_f(m, n, 1.0, -1, major, 0, 1)
        _, t = _time_calls(lib, hip, lambda: lib.ccfftmop_f(p, xv, yv), 20)
        y = np.empty((m, n), np.complex64)
        lib.cmcopyto_user_f(yv, 0, y.ctypes.data, None)
        vecs = y if major == 0 else y.T
        v = fftm.View(0, parent, m, 1, n)
        xs = fftm.gather(x, v)
        sel = f["rows"]
        want = fft64(xs[sel] if major == 0 else xs[:, sel].T, axis=1)
        ceil = size_ceiling(lib, 2 * m * n * 8)
        out[name] = {"us": round(t * 1e6, 2), "gb_per_s": round(2 * m * n * 8 / t / 1e9, 1),
                     "size_matched_ceiling_gb_per_s": ceil,
                     "frac_of_size_matched_ceiling": round(2 * m * n * 8 / t / 1e9 / ceil, 3) if ceil else None,
                     **col_pattern(lib, p, major, m, n, parent, t),
                     "rel_l2_vs_jvsip": rel_l2(vecs[sel], f["out_rows"]),
                     "rel_l2_vs_fp64": rel_l2(vecs[sel], want),
                     "schedule": json.loads(lib.describe(p))["passes"][0]["radix"]}
        lib.fftm_destroy_f(p)
        lib.cmalldestroy_f(yv)
        lib.cmdestroy_f(xv)
        lib.cmalldestroy_f(pv)
    # FP64: vsip_ccfftmip_d 2048 x 2048
    xd = datagen.complex_data(3, (2048, 2048), dtype=np.complex128)
    vd = lib.cmcreate_d(2048, 2048, 0, 0)
    lib.cmcopyfrom_user_d(np.ascontiguousarray(xd).ctypes.data, None, 0, vd)
    pd = lib.ccfftmip_create_d(2048, 2048, 1.0, -1, 0, 0, 1)
    pinv = lib.ccfftmip_create_d(2048, 2048, 1.0 / 2048, 1, 0, 0, 1)
    def fwdinv():
        lib.ccfftmip_d(pd, vd)
        lib.ccfftmip_d(pinv, vd)
    _, t = _time_calls(lib, hip, fwdinv, 10)
    lib.cmcopyfrom_user_d(np.ascontiguousarray(xd).ctypes.data, None, 0, vd)
    lib.ccfftmip_d(pd, vd)
    yd = np.empty((2048, 2048), np.complex128)
    lib.cmcopyto_user_d(vd, 0, yd.ctypes.data, None)
    rows = np.arange(0, 2048, 256)
    out["fp64_2048x2048"] = {"us_per_transform": round(t / 2 * 1e6, 2),
                             "gb_per_s": round(2 * 2048 * 2048 * 16 / (t / 2) / 1e9, 1),
                             "rel_l2_vs_numpy_fp64": rel_l2(yd[rows], fft64(xd[rows], axis=1))}
    lib.fftm_destroy_d(pd)
    lib.fftm_destroy_d(pinv)
    lib.cmalldestroy_d(vd)
    return out


def config_c4(lib, hip, njobs=3000, legs="all"):
    """njobs FFTs of log-uniform 2^a 3^b 5^c lengths in [64, 65536], rows so the
    total is ~2^27 complex elements: one batched graph launch vs plain per-call.
    legs="batched" times the graph replays only (for counter collection, where
    each of the plain legs' ~4000 launches would be serialised)"""
    from jvsip_amd import datagen, fftm
    lengths, rows = datagen.c4_jobs(njobs, 4)
    total = int(sum(r * n for r, n in zip(rows, lengths)))
    pool = datagen.complex_data(4, (total,))
    blk = lib.cblockcreate_f(total, 0)
    whole = lib.cmbind_f(blk, 0, total, 1, 1, total)
    lib.cmcopyfrom_user_f(pool.ctypes.data, None, 0, whole)
    ref_blk = lib.cblockcreate_f(total, 0)
    ref_whole = lib.cmbind_f(ref_blk, 0, total, 1, 1, total)
    views, rviews, plans, off = [], [], [], 0
    for r, n in zip(rows, lengths):
        views.append(lib.cmbind_f(blk, off, n, r, 1, n))
        rviews.append(lib.cmbind_f(ref_blk, off, n, r, 1, n))
        p = lib.ccfftmip_create_f(r, n, 1.0, -1, 0, 0, 1)
        assert p, lib.last_error()
        plans.append(p)
        off += r * n
    t0 = time.perf_counter()
    b = lib.batch(plans, views)
    t_create = time.perf_counter() - t0
    assert b, lib.last_error()
    info = lib.batch_info(b)
    # batched: graph replays (forward in place; the data just keeps transforming)
    wall_b, gpu_b = _time_calls(lib, hip, lambda: lib.amd_batch_run_f(b), 10)
    bytes_ = 2 * total * 8
    if legs == "batched":
        out = {"jobs": njobs, "elements": total, "distinct_lengths": len(set(lengths)), "batch": info,
               "batch_create_s": round(t_create, 3), "legs": legs,
               "batched_ms": round(gpu_b * 1e3, 3), "batched_gb_per_s": round(bytes_ / gpu_b / 1e9, 1)}
        lib.amd_batch_destroy_f(b)
        for p in plans:
            lib.fftm_destroy_f(p)
        for v in views + rviews:
            lib.cmdestroy_f(v)
        lib.cmalldestroy_f(whole)
        lib.cmalldestroy_f(ref_whole)
        return out
    # plain per-call, back to back (no synchronisation between calls)
    def plain_all():
        for p, v in zip(plans, views):
            lib.ccfftmip_f(p, v)
    wall_p, gpu_p = _time_calls(lib, hip, plain_all, 2)
    # plain per-call with a synchronisation after each call (sequential calls)
    t0 = time.perf_counter()
    for p, v in zip(plans, views):
        lib.ccfftmip_f(p, v)
        lib.amd_sync()
    wall_ps = time.perf_counter() - t0
    # accuracy: batched vs plain on the same fresh input, and a sample vs numpy's FP64 FFT
    lib.cmcopyfrom_user_f(pool.ctypes.data, None, 0, whole)
    lib.cmcopyfrom_user_f(pool.ctypes.data, None, 0, ref_whole)
    lib.amd_batch_run_f(b)
    for p, v in zip(plans, rviews):
        lib.ccfftmip_f(p, v)
    got = np.empty(total, np.complex64)
    want = np.empty(total, np.complex64)
    lib.cmcopyto_user_f(whole, 0, got.ctypes.data, None)
    lib.cmcopyto_user_f(ref_whole, 0, want.ctypes.data, None)
    errs, off = [], 0
    sample = set(range(0, njobs, max(1, njobs // 24)))
    worst_plain = 0.0
    for j, (r, n) in enumerate(zip(rows, lengths)):
        sl = slice(off, off + r * n)
        worst_plain = max(worst_plain, rel_l2(got[sl], want[sl]))
        if j in sample:
            errs.append(rel_l2(got[sl], fft64(pool[sl].reshape(r, n), axis=1)))
        off += r * n
    out = {"jobs": njobs, "elements": total, "distinct_lengths": len(set(lengths)), "batch": info,
           "batch_create_s": round(t_create, 3),
           "batched_ms": round(gpu_b * 1e3, 3), "batched_gb_per_s": round(bytes_ / gpu_b / 1e9, 1),
           "plain_back_to_back_ms": round(wall_p * 1e3, 3),
           "plain_sync_each_ms": round(wall_ps * 1e3, 3),
           "speedup_vs_plain_back_to_back": round(wall_p / wall_b, 2),
           "speedup_vs_plain_sync_each": round(wall_ps / wall_b, 2),
           "max_rel_l2_batched_vs_plain": worst_plain,
           "max_rel_l2_vs_fp64_sample": max(errs)}
    lib.amd_batch_destroy_f(b)
    for p in plans:
        lib.fftm_destroy_f(p)
    for v in views + rviews:
        lib.cmdestroy_f(v)
    lib.cmalldestroy_f(whole)
    lib.cmalldestroy_f(ref_whole)
    return out


def config_fft1d(lib, hip):
    """1-D vsip_ccfftip_f forward on single long vectors (the two-pass split
    over one vector) and vsip_rcfftop_f on a long real vector"""
    from jvsip_amd import datagen
    out = {}
    for logn in (16, 20, 24):
        n = 1 << logn
        x = datagen.complex_data(6, (n,), dist="uniform")
        v = lib.cvcreate_f(n, 0)
        lib.cvcopyfrom_user_f(x.ctypes.data, None, v)
        p = lib.ccfftip_create_f(n, 1.0, -1, 0, 1)
        assert v and p, lib.last_error()
        _, t = _time_calls(lib, hip, lambda: lib.ccfftip_f(p, v), 20)
        y = np.empty(n, np.complex64)
        lib.ccfftip_f(p, v)
        lib.cvcopyfrom_user_f(x.ctypes.data, None, v)
        lib.ccfftip_f(p, v)
        lib.cvcopyto_user_f(v, y.ctypes.data, None)
        want = np.fft.fft(x.astype(np.complex128))
        out["ccfftip_2^%d" % logn] = {"us": round(t * 1e6, 2), "gb_per_s": round(2 * n * 8 / t / 1e9, 1),
                                       "rel_l2_vs_numpy_fp64": float(np.linalg.norm(y - want) / np.linalg.norm(want)),
                                       "schedule": [q["n"] for q in json.loads(lib.describe(p))["passes"]]}
        lib.fft_destroy_f(p)
        lib.cvalldestroy_f(v)
    n = 1 << 22
    r = datagen.real_data(6, (n,))
    xv = lib.vcreate_f(n, 0)
    zv = lib.cvcreate_f(n // 2 + 1, 0)
    lib.vcopyfrom_user_f(r.ctypes.data, xv)
    p = lib.rcfftop_create_f(n, 1.0, 0, 1)
    _, t = _time_calls(lib, hip, lambda: lib.rcfftop_f(p, xv, zv), 20)
    z = np.empty(n // 2 + 1, np.complex64)
    lib.cvcopyto_user_f(zv, z.ctypes.data, None)
    want = np.fft.rfft(r.astype(np.float64))
    out["rcfftop_2^22"] = {"us": round(t * 1e6, 2), "gb_per_s": round((n * 4 + (n // 2 + 1) * 8) / t / 1e9, 1),
                           "rel_l2_vs_numpy_fp64": float(np.linalg.norm(z - want) / np.linalg.norm(want))}
    lib.fft_destroy_f(p)
    lib.valldestroy_f(xv)
    lib.cvalldestroy_f(zv)
    return out


def _corr_setup(lib, cplx, m, n, support=0, bias=1):
    """A vsip_[c]correlate1d_f object for reference length m and data length
    n on user blocks (N(0,1) data) admitted to the library; returns a dict
    with the call, the release that returns y, the cleanup and the inputs."""
    from jvsip_amd import datagen
    lag = m + n - 1 if support == 0 else n if support == 1 else n - m + 1
    if cplx:
        arrs = [datagen.complex_data(21, (m,)), datagen.complex_data(22, (n,)), np.zeros(lag, np.complex64)]
        blocks = [lib.cblockbind_f(a.ctypes.data, None, a.size, 0) for a in arrs]
        for b in blocks:
            lib.cblockadmit_f(b, 1)
        views = [lib.cvbind_f(b, 0, 1, a.size) for b, a in zip(blocks, arrs)]
        obj = lib.ccorr1d_create_f(m, n, support, 0, 1)
        run, destroy = lib.ccorrelate1d_f, lib.ccorr1d_destroy_f
    else:
        arrs = [datagen.real_data(21, (m,)), datagen.real_data(22, (n,)), np.zeros(lag, np.float32)]
        blocks = [lib.blockbind_f(a.ctypes.data, a.size, 0) for a in arrs]
        for b in blocks:
            lib.blockadmit_f(b, 1)
        views = [lib.vbind_f(b, 0, 1, a.size) for b, a in zip(blocks, arrs)]
        obj = lib.corr1d_create_f(m, n, support, 0, 1)
        run, destroy = lib.correlate1d_f, lib.corr1d_destroy_f
    assert obj and all(views), lib.last_error() if hasattr(lib, "last_error") else "create failed"

    def result():
        if cplx:
            r, i = ctypes.c_void_p(), ctypes.c_void_p()
            lib.cblockrelease_f(blocks[2], 1, ctypes.byref(r), ctypes.byref(i))
        else:
            lib.blockrelease_f(blocks[2], 1)
        return arrs[2].copy()

    def cleanup():
        destroy(obj)
        for v in views:
            (lib.cvdestroy_f if cplx else lib.vdestroy_f)(v)
        for b in blocks:
            (lib.cblockdestroy_f if cplx else lib.blockdestroy_f)(b)

    return {"call": lambda: run(obj, bias, *views), "result": result, "cleanup": cleanup,
            "h": arrs[0], "x": arrs[1], "lag": lag}


def _corr_numpy(h, x):
    """FULL unbiased correlation in float64 with numpy's FFT:
    y[i] = sum_j h[j] conj(x[i + j - (m - 1)]) / min(i + 1, m, m + n - 1 - i)."""
    m, n = h.size, x.size
    L = 1 << int(np.ceil(np.log2(m + n - 1)))
    c = np.fft.ifft(np.fft.fft(np.asarray(x, np.complex128), L) * np.conj(np.fft.fft(np.asarray(h, np.complex128), L)))
    i = np.arange(m + n - 1)
    y = np.conj(c[(i - (m - 1)) % L]) / np.minimum(np.minimum(i + 1, m), m + n - 1 - i)
    return y if np.iscomplexobj(h) else y.real


def _corr_length(mn):
    """The library's transform length for mn lags: the smallest 2^a 3^b 5^c
    >= mn, or the reference's power of two above mn when that is within
    1.15x of it."""
    p2 = 2
    while p2 <= mn:
        p2 *= 2
    smooth = min(v for v in (2 ** a * 3 ** b * 5 ** c for a in range(64) for b in range(41) for c in range(28)
                             if 2 ** a * 3 ** b * 5 ** c < p2) if v >= mn) if mn < p2 else p2
    return p2 if p2 <= 1.15 * smooth else smooth


def _conv_setup(lib, m, n, D=1, support=0):
    """A vsip_convolve1d_f object (kernel m, data n, no symmetry) on user
    blocks admitted to the library; the call, the release of y, the cleanup."""
    from jvsip_amd import datagen
    L = ((m + n - 2) // D + 1) if support == 0 else ((n - 1) // D + 1) if support == 1 else \
        ((n - 1) // D - (m - 1) // D + 1)
    arrs = [datagen.real_data(23, (m,)), datagen.real_data(24, (n,)), np.zeros(L, np.float32)]
    blocks = [lib.blockbind_f(a.ctypes.data, a.size, 0) for a in arrs]
    for b in blocks:
        lib.blockadmit_f(b, 1)
    views = [lib.vbind_f(b, 0, 1, a.size) for b, a in zip(blocks, arrs)]
    obj = lib.conv1d_create_f(views[0], 0, n, D, support, 0, 1)
    assert obj, "conv1d_create failed"

    def result():
        lib.blockrelease_f(blocks[2], 1)
        return arrs[2].copy()

    def cleanup():
        lib.conv1d_destroy_f(obj)
        for v in views:
            lib.vdestroy_f(v)
        for b in blocks:
            lib.blockdestroy_f(b)

    return {"call": lambda: lib.convolve1d_f(obj, views[1], views[2]), "result": result, "cleanup": cleanup,
            "h": arrs[0], "x": arrs[1], "len": L}


def config_corr1d(lib, hip):
    """vsip_ccorrelate1d_f and vsip_correlate1d_f, FULL support, unbiased,
    and vsip_convolve1d_f, FULL, no decimation: reference / kernel 1024, data
    2^20, on admitted unit-stride views; accuracy against numpy FP64"""
    out = {}
    m, n = 1024, 1 << 20
    for cplx in (True, False):
        s = _corr_setup(lib, cplx, m, n)
        _, t = _time_calls(lib, hip, s["call"], 20)
        y = s["result"]()
        want = _corr_numpy(s["h"], s["x"])
        esz = 8 if cplx else 4
        out[("ccorrelate1d" if cplx else "correlate1d") + "_f_1024x2^20"] = {
            "us": round(t * 1e6, 2), "transform_length": _corr_length(m + n - 1), "reference_length": 1 << 21,
            "io_bytes": (m + n + s["lag"]) * esz,
            "gb_per_s_io": round((m + n + s["lag"]) * esz / t / 1e9, 1),
            "rel_l2_vs_numpy_fp64": float(np.linalg.norm(y - want) / np.linalg.norm(want))}
        s["cleanup"]()
    # convolution, kernel 1024, data 2^20, FULL, no decimation
    s = _conv_setup(lib, m, n)
    _, t = _time_calls(lib, hip, s["call"], 20)
    y = s["result"]()
    L = 1 << int(np.ceil(np.log2(m + n - 1)))
    want = np.fft.irfft(np.fft.rfft(s["x"].astype(np.float64), L) * np.fft.rfft(s["h"].astype(np.float64), L),
                        L)[:m + n - 1]
    out["convolve1d_f_1024x2^20"] = {
        "us": round(t * 1e6, 2), "transform_length": _corr_length(m + n - 1),
        "io_bytes": (m + n + s["len"]) * 4, "gb_per_s_io": round((m + n + s["len"]) * 4 / t / 1e9, 1),
        "rel_l2_vs_numpy_fp64": float(np.linalg.norm(y - want) / np.linalg.norm(want))}
    s["cleanup"]()
    return out


def config_f3_komega(lib, hip):
    """k-omega beamformer step (the reference's kw.c komega()): the VSIPL call
    sequence (vmmul ROW, vmmul COL, rcfftmop ROW, ccfftmip COL, mcmagsq,
    smmul, madd: 7 launches) and the fused step (vsip_amd_komega_f: 2
    launches), at the reference's parameter file (128 sensors x 1024 samples)
    and at 1024 x 8192; plane waves plus uniform noise"""
    from jvsip_amd import komega as kw
    out = {}
    for m, n in ((128, 1024), (1024, 8192)):
        nf = n // 2 + 1
        x = kw.plane_waves(m, n, [(n // 8, 3, 1.0), (n // 5, m - 7, 0.7)], seed=7, noise=0.5)
        ks, kf = kw.KOmega(lib, m, n, 8), kw.KOmega(lib, m, n, 8)
        ds, df = kw.data_matrix(lib, x), kw.data_matrix(lib, x)
        ks.zero()
        kf.zero()
        ks.step(ds)
        kf.step_fused(df)
        g_s, g_f = ks.gram(), kf.gram()
        xt = x.astype(np.float64) * np.hanning(n + 2)[1:-1][None, :] * np.hanning(m + 2)[1:-1][:, None]
        y = np.fft.ifft(np.fft.rfft(xt, axis=1), axis=0) * m
        want = (y.real ** 2 + y.imag ** 2) / 8
        _, t_s = _time_calls(lib, hip, lambda: ks.step(ds), 50)
        _, t_f = _time_calls(lib, hip, lambda: kf.step_fused(df), 50)
        fused_bytes = 4 * m * n + 24 * m * nf
        out["%dx%d" % (m, n)] = {
            "sequence_us": round(t_s * 1e6, 2), "fused_us": round(t_f * 1e6, 2),
            "speedup": round(t_s / t_f, 2),
            "fused_gb_per_s": round(fused_bytes / t_f / 1e9, 1), "fused_algorithmic_bytes": fused_bytes,
            "rel_l2_fused_vs_sequence": float(np.linalg.norm(g_f - g_s) / np.linalg.norm(g_s)),
            "rel_l2_fused_vs_numpy_fp64": float(np.linalg.norm(g_f - want) / np.linalg.norm(want))}
        lib.malldestroy_f(ds)
        lib.malldestroy_f(df)
        ks.destroy()
        kf.destroy()
    return out


def config_c5(lib, nreq=2000, rate=20000.0):
    """Poisson requests at `rate`/s: rows U[1,512], lengths {256, 1000, 1024,
    2187, 4096, 16384}, fwd/inv 50/50, N(0,1) inputs from a device pool,
    64 GB HBM budget.  Three forms of the same stream: served (small
    requests on the resident kernel, the others over 4 launch lanes beside
    it), plain calls with a synchronisation after each,
    and plain calls back to back (no synchronisation; completion from an
    event per request).  Run at the workload's rate, at 5x that rate, and
    with every request present at once (saturation)."""
    from jvsip_amd import datagen
    pool_n = datagen.C5_POOL
    x = datagen.complex_data(5, (pool_n,))
    blk = lib.cblockcreate_f(pool_n, 0)
    pv = lib.cmbind_f(blk, 0, pool_n, 1, 1, pool_n)
    lib.cmcopyfrom_user_f(x.ctypes.data, None, 0, pv)
    reqs = datagen.c5_requests(nreq, rate, seed=5)
    budget = 64 << 30
    out = {"requests": nreq, "rate_per_s": rate, "hbm_budget_bytes": budget}

    def lat(res, stream, lanes=None):
        l = (res["t_done"] - np.array([q[0] for q in stream])) * 1e6
        out = {"p50_us": round(float(np.percentile(l, 50)), 1), "p99_us": round(float(np.percentile(l, 99)), 1),
               "mean_us": round(float(l.mean()), 1)}
        if lanes is not None:
            # the resident kernel's requests report stream == lanes
            for name, m in (("resident", res["stream"] == lanes), ("lanes", res["stream"] != lanes)):
                if m.any():
                    out[name + "_p50_us"] = round(float(np.percentile(l[m], 50)), 1)
                    out[name + "_p99_us"] = round(float(np.percentile(l[m], 99)), 1)
        return out

    def forms(stream):
        res_s, st_s = lib.serve(stream, blk, None, budget=budget, lanes=4, max_jobs=64)
        res_p, st_p = lib.serve(stream, blk, None, budget=budget, plain=True, sync_each=True)
        res_b, st_b = lib.serve(stream, blk, None, budget=budget, plain=True, sync_each=False)
        return {"served": dict(lat(res_s, stream, lanes=4), **st_s),
                "plain_sync_each": dict(lat(res_p, stream), **st_p),
                "plain_back_to_back": dict(lat(res_b, stream), **st_b),
                "served_vs_plain_sync_each": round(st_s["req_per_s"] / st_p["req_per_s"], 2),
                "served_vs_plain_back_to_back": round(st_s["req_per_s"] / st_b["req_per_s"], 2)}

    lib.serve(reqs[:50], blk, None, budget=1 << 30)            # warm the schedules
    out["at_rate"] = forms(reqs)
    fast = [(q[0] / 5.0,) + q[1:] for q in reqs]
    out["at_5x_rate"] = forms(fast)
    out["saturation"] = forms([(0.0,) + q[1:] for q in reqs])
    out["throughput_ratio_basis"] = ("saturation req/s of the served form over each plain form; "
                                     "at the workload's rate every form keeps up with arrivals")
    lib.cmalldestroy_f(pv)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the device reaches its sustained clocks only after some milliseconds of
    # work: the default warm-up runs ~25 ms of the same steps before timing
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-configs", action="store_true", help="skip the other workload configurations")
    ap.add_argument("--only", default="", help="comma-separated configuration names (or prefixes, "
                    "e.g. c3,c4) to run besides the headline; default all")
    ap.add_argument("--c4-legs", default="all", choices=["all", "batched"],
                    help="c4: 'batched' times the graph replays only (no plain per-call legs)")
    ap.add_argument("--settle-ms", type=float, default=60.0,
                    help="milliseconds of headline steps run before the warm-up steps, so the "
                         "device reaches its sustained clocks whatever --warmup is")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(each GPU would take its own slab of rows with no exchange step)" % args.gpus)
        sys.exit(2)

    from jvsip_amd import datagen
    from jvsip_amd.vsip import load
    lib = load()
    hip = Hip()
    stream = lib.amd_stream()

    x = datagen.complex_data(2, (M, N), dist="uniform").ravel()
    blocks, views = [], []
    for _ in range(NBUF):
        b = lib.cblockcreate_f(M * N, 0)
        assert b, lib.last_error()
        v = lib.cmbind_f(b, 0, N, M, 1, N)
        lib.cmcopyfrom_user_f(x.ctypes.data, None, 0, v)
        blocks.append(b)
        views.append(v)
    fwd = lib.ccfftmip_create_f(M, N, 1.0, -1, 0, 0, 1)
    inv = lib.ccfftmip_create_f(M, N, 1.0 / N, 1, 0, 0, 1)
    assert fwd and inv, lib.last_error()
    schedule = lib.describe(fwd)

    # step i: forward on matrix i % 4, inverse on matrix (i + 2) % 4; each
    # matrix is visited every second step, alternating forward and inverse
    def step(i):
        lib.ccfftmip_f(fwd, views[i % NBUF])
        lib.ccfftmip_f(inv, views[(i + NBUF // 2) % NBUF])

    # the device's own in-place read+write and read-only HBM rates (4 x 128 MiB),
    # measured first: they also start bringing the clocks up
    rmw, rd = ctypes.c_double(), ctypes.c_double()
    assert lib.amd_hbm_probe(BYTES_PER_CALL // 2, 20, ctypes.byref(rmw), ctypes.byref(rd)) == 0
    # clock settling: the device reaches its sustained clocks only after some
    # milliseconds under load; run the same steps for settle_ms, untimed
    ts = time.perf_counter()
    settle_steps = 0
    while (time.perf_counter() - ts) * 1e3 < args.settle_ms:
        for _ in range(8):
            step(settle_steps)
            settle_steps += 1
        assert lib.amd_sync() == 0, lib.last_error()
    settle_ms = (time.perf_counter() - ts) * 1e3
    for i in range(args.warmup):
        step(settle_steps + i)
    assert lib.amd_sync() == 0, lib.last_error()
    ev0, ev1 = hip.event(), hip.event()
    t0 = time.perf_counter()
    hip.record(ev0, stream)
    for i in range(args.steps):
        step(settle_steps + args.warmup + i)
    hip.record(ev1, stream)
    rc = lib.amd_sync()
    t1 = time.perf_counter()
    assert rc == 0, lib.last_error()
    err = lib.last_error()
    assert not err, err
    gpu_ms = hip.elapsed_ms(ev0, ev1)

    ms_step = (t1 - t0) * 1e3 / args.steps
    value = 2 * BYTES_PER_CALL / (ms_step * 1e-3) / 1e9
    # the timed region holds nothing but the 2 * steps kernel launches
    avg_launch_ms = gpu_ms / (2 * args.steps)
    achieved = BYTES_PER_CALL / (avg_launch_ms * 1e-3) / 1e9
    flops = 5.0 * N * np.log2(N) * M

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_ccfftmip_4096x4096.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    acc = accuracy(lib, None)
    for v in views:
        lib.cmalldestroy_f(v)
    lib.fftm_destroy_f(fwd)
    lib.fftm_destroy_f(inv)

    configs = None
    if not args.no_configs:
        configs = {}
        for name, fn in (("c1_plain_64x1024", lambda: config_c1(lib, hip)),
                         ("c2_host_buffers_4096x4096", lambda: config_host_buffers(lib)),
                         ("c2_real_8192x2048", lambda: config_c2_real(lib, hip)),
                         ("p2_rows", lambda: config_p2_rows(lib, hip)),
                         ("c3_strided_mixed_fp64", lambda: config_c3(lib, hip)),
                         ("c4_batched_uneven", lambda: config_c4(lib, hip, legs=args.c4_legs)),
                         ("c5_served", lambda: config_c5(lib)),
                         ("f1_fft1d", lambda: config_fft1d(lib, hip)),
                         ("corr1d", lambda: config_corr1d(lib, hip)),
                         ("f3_komega", lambda: config_f3_komega(lib, hip))):
            if args.only and not any(name.startswith(o.strip()) for o in args.only.split(",") if o.strip()):
                continue
            t0 = time.perf_counter()
            configs[name] = fn()
            configs[name]["bench_seconds"] = round(time.perf_counter() - t0, 2)
            print("bench: %s done in %.1f s" % (name, time.perf_counter() - t0), file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        share = cpu_share()
        threads = args.cpu_threads or share["cores"]
        cpu = cpu_baseline(threads)
        if cpu:
            cpu["host"] = share

    out = {
        "metric": "multi-FFT HBM GB/s (vsip_ccfftmip_f 4096x4096 c2c fwd+inv)",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_ms": round(settle_ms, 1),
        "settle_steps": settle_steps,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: iid uniform[-1,1] real/imag, splitmix64 seed 2",
        "config": {"workload": "vsip_ccfftmip_f ROW forward (scale 1) + inverse (scale 1/4096) "
                               "on 4096x4096 complex FP32, 4 device-resident matrices in rotation",
                   "rows": M, "n": N, "bytes_per_transform": BYTES_PER_CALL,
                   "parallelism": "single GPU"},
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "gflops_5nlog2n": round(2 * flops / (ms_step * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel_avg_us": round(avg_launch_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": BYTES_PER_CALL,
                     "measured_inplace_rmw_gb_per_s": round(rmw.value, 1),
                     "measured_read_only_gb_per_s": round(rd.value, 1),
                     "frac_of_measured_inplace_rmw": round(achieved / rmw.value, 4) if rmw.value else None},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "schedule": json.loads(schedule),
        "configs": configs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
