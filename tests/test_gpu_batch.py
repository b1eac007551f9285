"""GPU parity of the batched launch (vsip_amd_batch_*) through the C ABI.

Each job of a batch must give the same result as the plain per-call
vsip_ccfftmip_f / vsip_ccfftmop_f on the same data (within FP32 rounding of
the two schedules: rel L2 1e-6) and as the double-precision oracle (1e-5).
The captured graph is replayed several times, with ordinary library calls in
between, and every replay is checked.
"""
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW, VSIP_COL, VSIP_FFT_FWD, VSIP_FFT_INV
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _to_dev(lib, x, major=VSIP_ROW):
    m, n = x.shape
    v = lib.cmcreate_f(m, n, major, 0)
    assert v, lib.last_error()
    xx = np.ascontiguousarray(x, dtype=np.complex64)
    lib.cmcopyfrom_user_f(xx.ctypes.data, None, VSIP_ROW, v)
    return v


def _to_host(lib, v, m, n):
    out = np.empty((m, n), np.complex64)
    lib.cmcopyto_user_f(v, VSIP_ROW, out.ctypes.data, None)
    return out


JOBS = [  # (rows, length, major, direction, scale, out_of_place)
    (700, 64, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (300, 100, VSIP_ROW, VSIP_FFT_INV, 0.01, False),
    (90, 243, VSIP_ROW, VSIP_FFT_FWD, 1.0, True),
    (40, 1000, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (44, 1024, VSIP_ROW, VSIP_FFT_INV, 1.0 / 1024, False),
    (23, 1920, VSIP_ROW, VSIP_FFT_FWD, 2.0, True),
    (20, 2187, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (11, 4096, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (7, 6561, VSIP_ROW, VSIP_FFT_INV, 1.0, False),
    (5, 8192, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (4, 10000, VSIP_ROW, VSIP_FFT_FWD, 1.0, True),
    (3, 16384, VSIP_ROW, VSIP_FFT_INV, 0.5, False),
    (1, 65536, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (2, 3125, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (6, 1001, VSIP_ROW, VSIP_FFT_FWD, 1.0, False),
    (160, 96, VSIP_COL, VSIP_FFT_FWD, 1.0, False),   # 96 columns of length 160
    (1, 40000, VSIP_ROW, VSIP_FFT_FWD, 1.0, True),
]


def _plan(lib, m, n, major, d, s, op):
    create = lib.ccfftmop_create_f if op else lib.ccfftmip_create_f
    p = create(m, n, s, d, major, 0, 1)
    assert p, lib.last_error()
    return p


def test_batch_matches_plain_and_oracle(amd):
    lib = amd
    datas, ins, outs, plans = [], [], [], []
    for k, (m, n, major, d, s, op) in enumerate(JOBS):
        x = datagen.complex_data(100 + k, (m, n))
        datas.append(x)
        ins.append(_to_dev(lib, x))
        outs.append(_to_dev(lib, np.zeros((m, n), np.complex64)) if op else None)
        plans.append(_plan(lib, m, n, major, d, s, op))
    _case("batch create: %d jobs" % len(JOBS))
    b = lib.batch(plans, ins, [o if o else i for o, i in zip(outs, ins)])
    assert b, lib.last_error()
    info = lib.batch_info(b)
    print("  batch:", info, flush=True)
    assert info["two_pass_jobs"] >= 3

    # expected results: oracle and plain per-call runs on copies
    want, plain = [], []
    for k, (m, n, major, d, s, op) in enumerate(JOBS):
        v = fftm.dense_view(m, n)
        want.append(po.ccfftm(datas[k].ravel(), v, major=major, direction=d, scale=s).reshape(m, n))
        plain.append(fftm.ccfftm(lib, datas[k].ravel(), v, major=major, direction=d, scale=s).reshape(m, n))

    for rep in range(3):
        _case("batch run %d" % rep)
        for k, (m, n, major, d, s, op) in enumerate(JOBS):   # restore in-place inputs
            if not op and rep:
                xx = np.ascontiguousarray(datas[k])
                lib.cmcopyfrom_user_f(xx.ctypes.data, None, VSIP_ROW, ins[k])
        assert lib.amd_batch_run_f(b) == 0, lib.last_error()
        assert lib.amd_sync() == 0
        for k, (m, n, major, d, s, op) in enumerate(JOBS):
            got = _to_host(lib, outs[k] if op else ins[k], m, n)
            e_or = po.rel_l2(got, want[k])
            e_pl = po.rel_l2(got, plain[k])
            print("  job %2d %5dx%-6d major %d dir %+d: vs oracle %.2e vs plain %.2e"
                  % (k, m, n, major, d, e_or, e_pl), flush=True)
            assert e_or < 1e-5, (k, e_or)
            assert e_pl < 1e-6, (k, e_pl)
            if op:
                np.testing.assert_array_equal(_to_host(lib, ins[k], m, n), datas[k])
    assert not lib.last_error(), lib.last_error()
    lib.amd_batch_destroy_f(b)
    for p in plans:
        lib.fftm_destroy_f(p)
    for v in ins + [o for o in outs if o]:
        lib.cmalldestroy_f(v)


def test_tuner_reports_schedules(amd):
    for n in (64, 1000, 2187, 4096, 8192):
        t = amd.tune(n)
        print("  tune", n, t, flush=True)
        assert t["rc"] == 0 and t["threads"] == 512 and t["vpb"] >= 1
        assert int(np.prod(t["radix"])) == n
    assert amd.tune(16384)["rc"] == -2      # two passes in the batched shape
    assert amd.tune(17)["rc"] == -1         # no butterfly for 17


_SMOOTH = sorted({2 ** a * 3 ** b * 5 ** c * 7 ** d for a in range(17) for b in range(7) for c in range(5)
                  for d in range(3) if 8 <= 2 ** a * 3 ** b * 5 ** c * 7 ** d <= 65536})
_ELEVEN = [11, 13, 22, 26, 44, 52, 88, 143, 176, 208, 286, 352, 1001, 1144, 1408, 2002, 2816, 4004]


def _batch_draw(seed):
    """A random batch: (rows, length, major, direction, scale, out_of_place) per job."""
    rng = np.random.default_rng(7000 + seed)
    jobs = []
    total = 0
    for _ in range(int(rng.integers(8, 25))):
        c = rng.random()
        if c < 0.35:
            n = 1 << int(rng.integers(3, 17))
        elif c < 0.85:
            n = int(rng.choice(_SMOOTH))
        else:
            n = int(rng.choice(_ELEVEN))
        k = max(1, int(rng.integers(1, max(2, (1 << 17) // n + 1))))
        major = VSIP_ROW if rng.random() < 0.8 else VSIP_COL
        d = VSIP_FFT_FWD if rng.random() < 0.5 else VSIP_FFT_INV
        s = float(rng.choice([1.0, 0.5, 1.0 / n]))
        op = bool(rng.random() < 0.3)
        jobs.append((k, n, major, d, s, op) if major == VSIP_ROW else (n, k, major, d, s, op))
        total += n * k
        if total > (1 << 21):
            break
    return jobs


@pytest.mark.parametrize("seed", range(6))
def test_batch_sweep(amd, seed):
    """Random batches (8..24 jobs of random lengths, rows or columns, in and
    out of place, either direction): every job against the oracle (1e-5) and
    its plain call (1e-6), on two replays with the inputs restored between."""
    lib = amd
    jobs = _batch_draw(seed)
    for (m, n, major, d, s, op) in jobs:
        L = n if major == VSIP_ROW else m
        assert lib.tune(L)["rc"] in (0, -2), L        # a batched schedule, one or two passes
    datas, ins, outs, plans = [], [], [], []
    for k, (m, n, major, d, s, op) in enumerate(jobs):
        x = datagen.complex_data(200 + 31 * seed + k, (m, n))
        datas.append(x)
        ins.append(_to_dev(lib, x))
        outs.append(_to_dev(lib, np.zeros((m, n), np.complex64)) if op else None)
        plans.append(_plan(lib, m, n, major, d, s, op))
    _case("batch sweep seed %d: %d jobs %s" % (seed, len(jobs), [(m, n, major) for (m, n, major, _, _, _) in jobs]))
    b = lib.batch(plans, ins, [o if o else i for o, i in zip(outs, ins)])
    assert b, lib.last_error()
    print("  batch:", lib.batch_info(b), flush=True)
    want, plain = [], []
    for k, (m, n, major, d, s, op) in enumerate(jobs):
        v = fftm.dense_view(m, n)
        want.append(po.ccfftm(datas[k].ravel(), v, major=major, direction=d, scale=s).reshape(m, n))
        plain.append(fftm.ccfftm(lib, datas[k].ravel(), v, major=major, direction=d, scale=s).reshape(m, n))
    for rep in range(2):
        if rep:
            for k, (m, n, major, d, s, op) in enumerate(jobs):
                if not op:
                    xx = np.ascontiguousarray(datas[k])
                    lib.cmcopyfrom_user_f(xx.ctypes.data, None, VSIP_ROW, ins[k])
        assert lib.amd_batch_run_f(b) == 0, lib.last_error()
        assert lib.amd_sync() == 0
        worst = 0.0
        for k, (m, n, major, d, s, op) in enumerate(jobs):
            got = _to_host(lib, outs[k] if op else ins[k], m, n)
            e_or, e_pl = po.rel_l2(got, want[k]), po.rel_l2(got, plain[k])
            worst = max(worst, e_or)
            assert e_or < 1e-5, (seed, rep, k, m, n, major, e_or)
            assert e_pl < 1e-6, (seed, rep, k, m, n, major, e_pl)
            if op:
                np.testing.assert_array_equal(_to_host(lib, ins[k], m, n), datas[k])
        print("  replay %d: worst vs oracle %.2e" % (rep, worst), flush=True)
    assert not lib.last_error(), lib.last_error()
    lib.amd_batch_destroy_f(b)
    for p in plans:
        lib.fftm_destroy_f(p)
    for v in ins + [o for o in outs if o]:
        lib.cmalldestroy_f(v)
