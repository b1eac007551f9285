"""Bit-reproducible synthetic inputs.

A counter-based splitmix64 stream: draw i (i = 0, 1, ...) of seed s is
mix(s + (i + 1) * 0x9E3779B97F4A7C15) with the splitmix64 finaliser.  Only
integer arithmetic and exactly-representable float operations are used, so
the same seed gives the same bits on every machine:

  uniform(-1, 1):  2 * (z >> 11) * 2**-53 - 1, rounded to the target type
  normal(0, 1):    Irwin-Hall sum of 12 draws of (z >> 32), scaled by 2**-32,
                   minus 6 (the sum is an exact integer below 2**36)

Complex data take the real part from even draws and the imaginary part from
odd draws (per scalar: one draw for uniform, twelve consecutive for normal).
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, start, count):
    """Draws start .. start+count-1 of the stream `seed` as uint64."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, count, dtype=np.float32, start=0):
    z = splitmix64(seed, start, count)
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (2.0 * u - 1.0).astype(dtype)


def normal(seed, count, dtype=np.float32, start=0):
    z = splitmix64(seed, 12 * start, 12 * count)
    top = (z >> np.uint64(32)).reshape(count, 12)
    s = top.sum(axis=1, dtype=np.uint64)
    return (s.astype(np.float64) * 2.0 ** -32 - 6.0).astype(dtype)


def complex_data(seed, shape, dist="normal", dtype=np.complex64):
    """A complex array of `shape`, real/imag from even/odd scalar draws."""
    n = int(np.prod(shape))
    rdt = np.float32 if dtype == np.complex64 else np.float64
    draw = normal if dist == "normal" else uniform
    x = draw(seed, 2 * n, rdt)
    out = np.empty(n, dtype=dtype)
    out.real = x[0::2]
    out.imag = x[1::2]
    return out.reshape(shape)


def real_data(seed, shape, dist="normal", dtype=np.float32):
    n = int(np.prod(shape))
    draw = normal if dist == "normal" else uniform
    return draw(seed, n, dtype).reshape(shape)


# ---- the batched and served workloads ----------------------------------------

def c4_lengths(count, seed=4):
    """`count` lengths drawn log-uniform over the 2^a 3^b 5^c values in
    [64, 65536] (the nearest such value to each log-uniform draw)."""
    cands = sorted({2 ** a * 3 ** b * 5 ** c for a in range(17) for b in range(11) for c in range(7)
                    if 64 <= 2 ** a * 3 ** b * 5 ** c <= 65536})
    logc = np.log(np.array(cands, dtype=np.float64))
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(64), np.log(65536), count)
    return [cands[int(np.argmin(np.abs(logc - v)))] for v in u]


def c4_jobs(count=3000, seed=4, total=1 << 27):
    """(lengths, rows) of the uneven batch: each job's row count makes its
    share of the `total` complex elements about equal (at least one row)."""
    lengths = c4_lengths(count, seed)
    per_job = total / count
    rows = [max(1, int(round(per_job / n))) for n in lengths]
    return lengths, rows


C5_LENGTHS = (256, 1000, 1024, 2187, 4096, 16384)
C5_POOL = 512 * 16384 * 2        # complex elements of the device input pool


def c5_requests(count=2000, rate=20000.0, seed=5, pool=C5_POOL, lengths=C5_LENGTHS, max_rows=512):
    """A Poisson request stream at `rate` per second: rows uniform in
    [1, max_rows], length uniform over `lengths`, forward (scale 1) or inverse
    (scale 1/n) 50/50, input at a uniform offset of the input pool; tuples in
    vsip_amd_request field order with out_offset -1 (output from the
    server's budget)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    reqs = []
    for i in range(count):
        n = int(rng.choice(lengths))
        r = int(rng.integers(1, max_rows + 1))
        d = -1 if rng.random() < 0.5 else 1
        reqs.append((float(t[i]), r, n, d, 1.0 if d < 0 else 1.0 / n, int(rng.integers(0, pool - r * n)), -1))
    return reqs
