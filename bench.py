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
                and against the double-precision oracle
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
    return {
        "value": round(gbs_all, 4), "unit": "GB/s", "cores": threads, "kind": "reference",
        "sample": "jvsip vsip_ccfftmip_f fwd+inv on the full 4096x4096 matrix as %d row slabs "
                  "(one plan per thread), %d timed steps after 1 warm-up; %.3f s per step" % (threads, steps, dt),
        "single_core": {"value": round(gbs_one, 4), "unit": "GB/s", "cores": 1,
                        "sample": "fwd+inv on %d rows x 4096, %.3f s" % (rows_single, dt1)},
        "cpu_model": cpu,
    }


def accuracy(lib, view_fn):
    """Forward transform of the seed-2 input: 8 rows against jvsip and the oracle."""
    from jvsip_amd import datagen, fftm
    from oracle import pyoracle as po
    f = np.load(os.path.join(ROOT, "tests", "golden", "jvsip_c2_4096x4096_fwd.npz"))
    x = datagen.complex_data(2, (M, N), dist="uniform")
    y = fftm.ccfftm(lib, x.ravel(), fftm.dense_view(M, N)).reshape(M, N)
    rows = f["rows"]
    want = po.ccfftm(x[rows].ravel(), fftm.dense_view(len(rows), N)).reshape(len(rows), N)
    return {"rel_l2_vs_jvsip": po.rel_l2(y[rows], f["out_rows"]),
            "rel_l2_vs_fp64_oracle": po.rel_l2(y[rows], want),
            "rows": [int(r) for r in rows]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the multiple FFT shards by rows with no exchange, see DESIGN.md)" % args.gpus)
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

    for i in range(args.warmup):
        step(i)
    assert lib.amd_sync() == 0, lib.last_error()
    ev0, ev1 = hip.event(), hip.event()
    t0 = time.perf_counter()
    hip.record(ev0, stream)
    for i in range(args.steps):
        step(args.warmup + i)
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

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(threads)

    out = {
        "metric": "multi-FFT HBM GB/s (vsip_ccfftmip_f 4096x4096 c2c fwd+inv)",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
                     "algorithmic_bytes_per_launch": BYTES_PER_CALL},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "schedule": json.loads(schedule),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
