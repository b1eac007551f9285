"""GPU tests of the continuous-batching server (vsip_amd_serve_f) through the C ABI.

A request stream with the serving workload's lengths {256, 1000, 1024, 2187,
4096, 16384}, mixed directions and scales arrives over time; every request
reads its input from a device pool and writes its own region of an output
pool.  Each result must be the same when the request is served alone as in
the mix (rel L2 <= 1e-6; the same schedule runs it, so in practice the bits
agree), agree with the plain sequential calls (1e-6) and with the
double-precision oracle (1e-5).  A tight HBM budget forces admission to wait
for memory, and must still complete every request with the same results.
"""
import sys

import numpy as np
import pytest

from jvsip_amd import datagen, fftm
from jvsip_amd.vsip import VSIP_ROW
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

LENGTHS = [256, 1000, 1024, 2187, 4096, 16384]
POOL = 1 << 21            # complex elements in the input pool


def _case(name):
    print("\n[case] %s" % name, flush=True)
    sys.stdout.flush()


def _requests(nreq, rate, seed, max_rows):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs, out_off = [], 0
    for i in range(nreq):
        n = int(rng.choice(LENGTHS))
        rows = int(rng.integers(1, max_rows + 1))
        d = -1 if rng.random() < 0.5 else 1
        scale = 1.0 if d < 0 else 1.0 / n
        in_off = int(rng.integers(0, POOL - rows * n))
        reqs.append((float(t[i]), rows, n, d, scale, in_off, out_off))
        out_off += rows * n
    return reqs, out_off


@pytest.fixture(scope="module")
def pools(amd):
    lib = amd
    x = datagen.complex_data(5, (POOL,))
    inb = lib.cblockcreate_f(POOL, 0)
    inv = lib.cmbind_f(inb, 0, POOL, 1, 1, POOL)
    lib.cmcopyfrom_user_f(x.ctypes.data, None, VSIP_ROW, inv)
    yield x, inb
    lib.cmalldestroy_f(inv)


def _out_pool(lib, total):
    b = lib.cblockcreate_f(total, 0)
    v = lib.cmbind_f(b, 0, total, 1, 1, total)
    return b, v


def _read(lib, v, total):
    out = np.empty(total, np.complex64)
    lib.cmcopyto_user_f(v, VSIP_ROW, out.ctypes.data, None)
    return out


def test_serve_mix_equals_alone_plain_and_oracle(amd, pools):
    lib = amd
    x, inb = pools
    reqs, total = _requests(48, rate=20000.0, seed=5, max_rows=48)
    ob, ov = _out_pool(lib, total)
    _case("serve mix: %d requests" % len(reqs))
    res, stats = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=4, max_jobs=16)
    print("  stats", stats, flush=True)
    assert np.all(res["t_done"] >= res["t_admit"]) and np.all(res["t_admit"] >= 0)
    mix = _read(lib, ov, total)

    _case("serve each request alone")
    alone = np.empty_like(mix)
    for r in reqs:
        lib.serve([(0.0,) + tuple(r[1:])], inb, ob, budget=1 << 28, lanes=1, max_jobs=1)
    alone[:] = _read(lib, ov, total)

    _case("plain sequential calls")
    res_p, stats_p = lib.serve(reqs, inb, ob, budget=1 << 30, plain=True)
    print("  plain stats", stats_p, flush=True)
    plain = _read(lib, ov, total)

    for i, (t, rows, n, d, s, io, oo) in enumerate(reqs):
        sl = slice(oo, oo + rows * n)
        want = po.ccfftm(x[io:io + rows * n], fftm.dense_view(rows, n), direction=d, scale=s)
        e_alone = po.rel_l2(mix[sl], alone[sl])
        e_plain = po.rel_l2(mix[sl], plain[sl])
        e_or = po.rel_l2(mix[sl], want)
        print("  req %2d %4dx%-6d dir %+d: alone %.1e plain %.1e oracle %.1e"
              % (i, rows, n, d, e_alone, e_plain, e_or), flush=True)
        assert e_alone <= 1e-6 and e_plain <= 1e-6 and e_or <= 1e-5, i
    lib.cmalldestroy_f(ov)


def test_serve_budget_admission(amd, pools):
    """A budget smaller than the stream's in-flight demand: admission waits for
    memory, every request still completes with the right result."""
    lib = amd
    x, inb = pools
    reqs, _ = _requests(40, rate=1e6, seed=7, max_rows=32)
    reqs = [r[:6] + (-1,) for r in reqs]          # outputs from the server's budget
    biggest = max(2 * r[1] * r[2] * 8 for r in reqs)
    budget = 2 * biggest
    _case("serve with a %d-byte budget" % budget)
    res, stats = lib.serve(reqs, inb, None, budget=budget, lanes=4, max_jobs=8)
    print("  stats", stats, flush=True)
    assert stats["hbm_peak"] <= budget
    assert np.all(res["t_done"] > 0)
    assert stats["batches"] > 1


def test_serve_error_paths(amd, pools):
    lib = amd
    x, inb = pools
    with pytest.raises(RuntimeError):   # a request that can never fit the budget
        lib.serve([(0.0, 4, 4096, -1, 1.0, 0, -1)], inb, None, budget=1 << 16)
    with pytest.raises(RuntimeError):   # no butterfly for 17
        lib.serve([(0.0, 4, 17, -1, 1.0, 0, -1)], inb, None, budget=1 << 26)
    lib.amd_clear_error()   # the library keeps its last error; later tests check it


def _requests_two_pass(nreq, seed):
    """lengths that split into two passes (32768, 65536, 40000 = 2^6 5^4) and
    single-pass ones, many present at once, on a small input pool"""
    rng = np.random.default_rng(seed)
    reqs, out_off = [], 0
    for i in range(nreq):
        n = int(rng.choice([32768, 65536, 40000, 4096]))
        rows = int(rng.integers(1, 9))
        d = -1 if rng.random() < 0.5 else 1
        scale = 1.0 if d < 0 else 1.0 / n
        reqs.append((1e-6 * i, rows, n, d, scale, int(rng.integers(0, POOL - rows * n)), out_off))
        out_off += rows * n
    return reqs, out_off


@pytest.mark.parametrize("direct_elems", [None, "1000000000"])
def test_serve_two_pass_lengths_on_four_lanes(amd, pools, monkeypatch, direct_elems):
    """Requests of one two-pass length run on several lanes at once, each on
    its own scratch: the mix equals each request alone, the plain calls and
    the oracle.  With the direct-launch threshold raised, the small ones run
    in the lanes' batched launches (their two-pass form)."""
    lib = amd
    x, inb = pools
    if direct_elems:
        monkeypatch.setenv("VSIP_AMD_SERVE_DIRECT", direct_elems)
    reqs, total = _requests_two_pass(32, seed=11)
    ob, ov = _out_pool(lib, total)
    _case("serve two-pass lengths, 4 lanes, direct threshold %s" % (direct_elems or "default"))
    res, stats = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=4, max_jobs=4)
    print("  stats", stats, flush=True)
    assert stats["max_lanes_in_flight"] > 1
    if direct_elems:
        assert stats["batched_requests"] > 0
    mix = _read(lib, ov, total)
    res_a, _ = lib.serve([(0.0,) + r[1:] for r in reqs], inb, ob, budget=1 << 30, lanes=1, max_jobs=1)
    alone = _read(lib, ov, total)
    lib.serve(reqs, inb, ob, budget=1 << 30, plain=True)
    plain = _read(lib, ov, total)
    for i, (t, rows, n, d, s, io, oo) in enumerate(reqs):
        sl = slice(oo, oo + rows * n)
        want = po.ccfftm(x[io:io + rows * n], fftm.dense_view(rows, n), direction=d, scale=s)
        e_alone, e_plain, e_or = po.rel_l2(mix[sl], alone[sl]), po.rel_l2(mix[sl], plain[sl]), po.rel_l2(mix[sl], want)
        print("  req %2d %dx%-6d dir %+d: alone %.1e plain %.1e oracle %.1e" % (i, rows, n, d, e_alone, e_plain, e_or),
              flush=True)
        assert e_alone <= 1e-6 and e_plain <= 1e-6 and e_or <= 1e-5, i
    lib.cmalldestroy_f(ov)


def test_serve_request_beyond_a_lane_runs_directly(amd, monkeypatch):
    """A request whose batched work items exceed one lane's item table
    (max_jobs * 1200; one 8000-point row per item here) launches its own
    kernel instead of waiting for a lane that could never hold it."""
    lib = amd
    monkeypatch.setenv("VSIP_AMD_SERVE_DIRECT", "1000000000")
    rows, n = 1300, 8000
    big = rows * n + 3000
    x = datagen.complex_data(15, (big,))
    inb = lib.cblockcreate_f(big, 0)
    inv = lib.cmbind_f(inb, 0, big, 1, 1, big)
    lib.cmcopyfrom_user_f(x.ctypes.data, None, VSIP_ROW, inv)
    reqs = [(0.0, rows, n, -1, 1.0, 0, 0), (0.0, 3, 1000, 1, 1e-3, 7, rows * n)]
    ob, ov = _out_pool(lib, big)
    _case("serve %dx%d with max_jobs 1" % (rows, n))
    res, stats = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=2, max_jobs=1)
    print("  stats", stats, flush=True)
    assert stats["direct_launched"] >= 1
    mix = _read(lib, ov, big)
    for t, r, nn, d, s, io, oo in reqs:
        want = po.ccfftm(x[io:io + r * nn], fftm.dense_view(r, nn), direction=d, scale=s)
        assert po.rel_l2(mix[oo:oo + r * nn], want) <= 1e-5
    lib.cmalldestroy_f(ov)
    lib.cmalldestroy_f(inv)


def _small_requests(nreq, rate, seed):
    """requests of the serving workload's lengths small enough for the
    resident kernel (rows * n <= 131072)"""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs, out_off = [], 0
    for i in range(nreq):
        n = int(rng.choice([256, 1000, 1024, 2187, 4096, 6000]))
        rows = int(rng.integers(1, min(64, 131072 // n) + 1))
        d = -1 if rng.random() < 0.5 else 1
        scale = 1.0 if d < 0 else 1.0 / n
        reqs.append((float(t[i]), rows, n, d, scale, int(rng.integers(0, POOL - rows * n)), out_off))
        out_off += rows * n
    return reqs, out_off


@pytest.mark.parametrize("rate", [2000.0, 1e6])
def test_serve_resident_mix_equals_alone_plain_and_oracle(amd, pools, rate):
    """Every request small enough runs on the resident kernel, which stays on
    the device for the stream: at 2000 req/s its workgroups wait between
    arrivals (the host publishes each request while the kernel runs); at 1e6
    req/s many requests are in flight on it at once.  Each result equals the
    request served alone (a one-request stream on the same path), the plain
    calls and the oracle."""
    lib = amd
    x, inb = pools
    reqs, total = _small_requests(40, rate, seed=21)
    ob, ov = _out_pool(lib, total)
    _case("resident mix at %g req/s: %d requests" % (rate, len(reqs)))
    res, stats = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=4, max_jobs=16)
    print("  stats", stats, flush=True)
    assert stats["resident_requests"] == len(reqs) and stats["direct_launched"] == 0
    assert np.all(res["stream"] == 4) and np.all(res["t_done"] >= res["t_admit"])
    mix = _read(lib, ov, total)
    _case("each request alone")
    for r in reqs:
        _, st1 = lib.serve([(0.0,) + tuple(r[1:])], inb, ob, budget=1 << 28, lanes=1, max_jobs=1)
        assert st1["resident_requests"] == 1
    alone = _read(lib, ov, total)
    lib.serve(reqs, inb, ob, budget=1 << 30, plain=True)
    plain = _read(lib, ov, total)
    for i, (t, rows, n, d, s, io, oo) in enumerate(reqs):
        sl = slice(oo, oo + rows * n)
        want = po.ccfftm(x[io:io + rows * n], fftm.dense_view(rows, n), direction=d, scale=s)
        e_alone, e_plain, e_or = po.rel_l2(mix[sl], alone[sl]), po.rel_l2(mix[sl], plain[sl]), po.rel_l2(mix[sl], want)
        print("  req %2d %3dx%-5d dir %+d: alone %.1e plain %.1e oracle %.1e" % (i, rows, n, d, e_alone, e_plain, e_or),
              flush=True)
        assert e_alone == 0.0 and e_plain <= 1e-6 and e_or <= 1e-5, i
    lib.cmalldestroy_f(ov)


def test_serve_resident_and_lanes_together(amd, pools, monkeypatch):
    """Small requests on the resident kernel while large ones (and two-pass
    lengths) launch on the lanes beside it, with outputs from the budget; the
    same stream with the resident path off gives the same results."""
    lib = amd
    x, inb = pools
    reqs, total = _requests(48, rate=50000.0, seed=23, max_rows=48)
    reqs += [(reqs[-1][0] + 1e-5 * (k + 1), 3, 32768, -1, 1.0, 11 * k, total + 3 * 32768 * k) for k in range(4)]
    total += 4 * 3 * 32768
    ob, ov = _out_pool(lib, total)
    _case("resident and lanes together: %d requests" % len(reqs))
    _, st = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=4, max_jobs=16)
    print("  stats", st, flush=True)
    assert st["resident_requests"] > 0 and st["direct_launched"] > 0
    both = _read(lib, ov, total)
    monkeypatch.setenv("VSIP_AMD_SERVE_RESIDENT", "0")
    _, st0 = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=4, max_jobs=16)
    assert st0["resident_requests"] == 0
    off = _read(lib, ov, total)
    for i, (t, rows, n, d, s, io, oo) in enumerate(reqs):
        sl = slice(oo, oo + rows * n)
        want = po.ccfftm(x[io:io + rows * n], fftm.dense_view(rows, n), direction=d, scale=s)
        assert po.rel_l2(both[sl], off[sl]) <= 1e-6 and po.rel_l2(both[sl], want) <= 1e-5, i
    lib.cmalldestroy_f(ov)


def test_serve_resident_spin_limit_stops_cleanly(amd, pools, monkeypatch):
    """A waiting resident workgroup gives up after VSIP_AMD_SERVE_SPIN_LIMIT
    polls: with a limit of 1 and requests 2 ms apart the kernel stops before
    the stream is over, the server reports an error instead of waiting for
    it, and the next stream runs normally."""
    lib = amd
    x, inb = pools
    reqs, total = _small_requests(6, 500.0, seed=25)
    ob, ov = _out_pool(lib, total)
    monkeypatch.setenv("VSIP_AMD_SERVE_SPIN_LIMIT", "1")
    _case("resident kernel with a spin limit of 1")
    with pytest.raises(RuntimeError, match="resident"):
        lib.serve(reqs, inb, ob, budget=1 << 30, lanes=2, max_jobs=4)
    lib.amd_clear_error()
    monkeypatch.delenv("VSIP_AMD_SERVE_SPIN_LIMIT")
    _, st = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=2, max_jobs=4)
    assert st["resident_requests"] == len(reqs)
    got = _read(lib, ov, total)
    for t, rows, n, d, s, io, oo in reqs:
        want = po.ccfftm(x[io:io + rows * n], fftm.dense_view(rows, n), direction=d, scale=s)
        assert po.rel_l2(got[oo:oo + rows * n], want) <= 1e-5
    lib.cmalldestroy_f(ov)


def test_serve_plain_back_to_back(amd, pools):
    """The plain baseline without a synchronisation per call gives the same
    results and completion times after the arrivals."""
    lib = amd
    x, inb = pools
    reqs, total = _requests(40, rate=1e6, seed=9, max_rows=32)
    ob, ov = _out_pool(lib, total)
    _case("plain calls back to back")
    res, stats = lib.serve(reqs, inb, ob, budget=1 << 30, plain=True, sync_each=False)
    print("  stats", stats, flush=True)
    assert stats["sync_each"] == 0
    assert np.all(res["t_done"] >= res["t_admit"]) and np.all(np.diff(res["t_done"]) >= 0)
    got = _read(lib, ov, total)
    for t, rows, n, d, s, io, oo in reqs:
        want = po.ccfftm(x[io:io + rows * n], fftm.dense_view(rows, n), direction=d, scale=s)
        assert po.rel_l2(got[oo:oo + rows * n], want) <= 1e-5
    lib.cmalldestroy_f(ov)


_SWEEP_LENGTHS = [256, 512, 1000, 1001, 1024, 2187, 3000, 4096, 6000, 8192, 16384, 32768, 40000, 65536]


@pytest.mark.parametrize("seed", range(6))
def test_serve_sweep(amd, pools, seed):
    """Random request streams (20..80 requests of mixed lengths, including
    two-pass ones, 1..64 rows, arrival rates from spread out to all at once,
    1..4 lanes): every served result equals the plain calls' (1e-6) and a
    sample the oracle's (1e-5)."""
    lib = amd
    x, inb = pools
    rng = np.random.default_rng(9000 + seed)
    rate = float(rng.choice([1e3, 2e4, 1e6]))
    t = np.cumsum(rng.exponential(1.0 / rate, int(rng.integers(20, 81))))
    reqs, total = [], 0
    for ti in t:
        n = int(rng.choice(_SWEEP_LENGTHS))
        rows = int(rng.integers(1, min(64, POOL // n) + 1))
        d = -1 if rng.random() < 0.5 else 1
        scale = float(rng.choice([1.0, 0.5, 1.0 / n]))
        reqs.append((float(ti), rows, n, d, scale, int(rng.integers(0, POOL - rows * n + 1)), total))
        total += rows * n
    lanes, max_jobs = int(rng.integers(1, 5)), int(rng.choice([8, 16, 64]))
    _case("serve sweep seed %d: %d requests at %g/s, %d lanes, %d jobs" % (seed, len(reqs), rate, lanes, max_jobs))
    ob, ov = _out_pool(lib, total)
    res, stats = lib.serve(reqs, inb, ob, budget=1 << 30, lanes=lanes, max_jobs=max_jobs)
    print("  stats", stats, flush=True)
    served = _read(lib, ov, total)
    lib.serve(reqs, inb, ob, budget=1 << 30, plain=True, sync_each=False)
    plain = _read(lib, ov, total)
    for i, (ti, rows, n, d, s, io, oo) in enumerate(reqs):
        sl = slice(oo, oo + rows * n)
        e_pl = po.rel_l2(served[sl], plain[sl])
        assert e_pl <= 1e-6, (seed, i, rows, n, e_pl)
        if i % 7 == 0:
            want = po.ccfftm(x[io:io + rows * n], fftm.dense_view(rows, n), direction=d, scale=s)
            e_or = po.rel_l2(served[sl], want)
            assert e_or <= 1e-5, (seed, i, rows, n, e_or)
    assert np.all(res["t_done"] >= res["t_admit"])
    lib.cmalldestroy_f(ov)
