"""vsip_amd_batch_run_noalloc_f (the batched launch that allocates nothing)
against vsip_amd_batch_create_f + vsip_amd_batch_run_f, through the C ABI.

Two calls in a row, the second with a smaller input (fewer jobs, fewer
rows), issued back to back before one synchronisation; each call's outputs
against the created batch's on the same inputs (rel L2 <= 1e-6 per job; the
same kernel and schedules run, so the bits agree in practice) and a sample
against the oracle (<= 1e-5).  The entry point refuses, before any launch,
lengths that are not tuned or not single pass in the batched shape and
inputs beyond its scratch.
"""
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# (rows, length, out of place): powers of two take the batched power-of-two
# stages, the rest the generic stages
JOBS = [(40, 256, 0), (3, 8192, 0), (17, 1000, 1), (9, 2187, 0), (64, 512, 1), (5, 1920, 0), (2, 4096, 1),
        (31, 100, 0), (7, 6561, 0), (12, 3000, 1), (1, 64, 0), (130, 1024, 0), (4, 5000, 1), (11, 243, 0)]


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


class _Jobs:
    """plans and views of one set of jobs over two device blocks (inputs, and
    outputs for the out-of-place jobs), filled from a seeded pool"""

    def __init__(self, lib, jobs, seed):
        self.lib, self.jobs = lib, jobs
        self.total = sum(r * n for r, n, _ in jobs)
        self.x = datagen.complex_data(seed, (self.total,))
        self.blk = lib.cblockcreate_f(self.total, 0)
        self.oblk = lib.cblockcreate_f(self.total, 0)
        self.whole = lib.cmbind_f(self.blk, 0, self.total, 1, 1, self.total)
        self.owhole = lib.cmbind_f(self.oblk, 0, self.total, 1, 1, self.total)
        assert self.blk and self.oblk and self.whole and self.owhole, lib.last_error()
        self.plans, self.ins, self.outs, self.offs = [], [], [], []
        off = 0
        for r, n, op in jobs:
            self.ins.append(lib.cmbind_f(self.blk, off, n, r, 1, n))
            self.outs.append(lib.cmbind_f(self.oblk, off, n, r, 1, n) if op else None)
            create = lib.ccfftmop_create_f if op else lib.ccfftmip_create_f
            p = create(r, n, 1.0 if op else 0.5, -1 if op else 1, VSIP_ROW, 0, 1)
            assert p, lib.last_error()
            self.plans.append(p)
            self.offs.append(off)
            off += r * n

    def reset(self):
        self.lib.cmcopyfrom_user_f(self.x.ctypes.data, None, VSIP_ROW, self.whole)
        z = np.zeros(self.total, np.complex64)
        self.lib.cmcopyfrom_user_f(z.ctypes.data, None, VSIP_ROW, self.owhole)

    def result(self):
        a = np.empty(self.total, np.complex64)
        b = np.empty(self.total, np.complex64)
        self.lib.cmcopyto_user_f(self.whole, VSIP_ROW, a.ctypes.data, None)
        self.lib.cmcopyto_user_f(self.owhole, VSIP_ROW, b.ctypes.data, None)
        return [(b if op else a)[o:o + r * n] for (r, n, op), o in zip(self.jobs, self.offs)]

    def outs_arg(self):
        return [o if o else v for o, v in zip(self.outs, self.ins)]

    def destroy(self):
        for p in self.plans:
            self.lib.fftm_destroy_f(p)
        for v in self.ins + [o for o in self.outs if o]:
            self.lib.cmdestroy_f(v)
        self.lib.cmalldestroy_f(self.whole)
        self.lib.cmalldestroy_f(self.owhole)


def _created_batch(lib, js):
    js.reset()
    b = lib.batch(js.plans, js.ins, js.outs_arg())
    assert b, lib.last_error()
    assert lib.amd_batch_run_f(b) == 0, lib.last_error()
    out = js.result()
    lib.amd_batch_destroy_f(b)
    return out


def test_noalloc_matches_created_batch(amd):
    lib = amd
    lib.amd_clear_error()
    for r, n, op in JOBS:
        t = lib.tune(n)
        assert t["rc"] == 0, (n, t)
    big = _Jobs(lib, JOBS, 41)
    small_jobs = [(max(1, r // 3), n, op) for r, n, op in JOBS[3:10]]
    small = _Jobs(lib, small_jobs, 42)

    _case("created batch: %d jobs, then %d jobs" % (len(JOBS), len(small_jobs)))
    want_big = _created_batch(lib, big)
    want_small = _created_batch(lib, small)

    _case("no-alloc: two calls back to back (the second smaller), one synchronisation")
    big.reset()
    small.reset()
    used0 = lib.amd_bytes_in_use()
    assert lib.batch_noalloc(big.plans, big.ins, big.outs_arg()) == 0, lib.last_error()
    assert lib.batch_noalloc(small.plans, small.ins, small.outs_arg()) == 0, lib.last_error()
    assert lib.amd_sync() == 0
    assert lib.amd_bytes_in_use() == used0, "the no-alloc calls allocated device memory"
    got_big, got_small = big.result(), small.result()
    for name, js, got, want in (("first", big, got_big, want_big), ("second", small, got_small, want_small)):
        for j, ((r, n, op), g, w) in enumerate(zip(js.jobs, got, want)):
            e = po.rel_l2(g, w)
            print("  %s call job %2d %4d x %-5d: vs created batch %.1e" % (name, j, r, n, e), flush=True)
            assert e <= 1e-6, (name, j)
        for j in (0, 2, len(js.jobs) - 1):
            r, n, op = js.jobs[j]
            o = js.offs[j]
            want_o = po.ccfftm(js.x[o:o + r * n], fftm.dense_view(r, n), direction=-1 if op else 1,
                               scale=1.0 if op else 0.5)
            e = po.rel_l2(got[j], want_o)
            print("  %s call job %2d vs oracle %.1e" % (name, j, e), flush=True)
            assert e <= 1e-5

    _case("no-alloc: the same two calls again, each synchronised")
    for js, want in ((big, want_big), (small, want_small)):
        js.reset()
        assert lib.batch_noalloc(js.plans, js.ins, js.outs_arg()) == 0, lib.last_error()
        assert lib.amd_sync() == 0
        for g, w in zip(js.result(), want):
            assert po.rel_l2(g, w) <= 1e-6
    assert not lib.last_error(), lib.last_error()
    big.destroy()
    small.destroy()


def test_noalloc_refuses_what_it_cannot_hold(amd):
    lib = amd
    x = _Jobs(lib, [(2, 1210, 0)], 43)              # 1210 = 2 5 11^2, never tuned here
    lib.amd_clear_error()
    assert lib.batch_noalloc(x.plans, x.ins) == -1
    assert "tune" in lib.last_error()
    x.destroy()
    assert lib.tune(16384)["rc"] == -2              # two passes in the batched shape
    y = _Jobs(lib, [(1, 16384, 0)], 44)
    lib.amd_clear_error()
    assert lib.batch_noalloc(y.plans, y.ins) == -1
    assert "single pass" in lib.last_error()
    y.destroy()
    # sizes beyond the scratch (one plan and view repeated: nothing runs)
    w = _Jobs(lib, [(130, 1024, 0)], 45)
    assert lib.tune(1024)["rc"] == 0
    lib.amd_clear_error()
    assert lib.batch_noalloc(w.plans * 10000, w.ins * 10000) == -1      # descriptors
    assert "scratch" in lib.last_error()
    lib.amd_clear_error()
    assert lib.batch_noalloc(w.plans * 8000, w.ins * 8000) == -1        # work items
    assert "scratch" in lib.last_error()
    w.destroy()
    lib.amd_clear_error()
