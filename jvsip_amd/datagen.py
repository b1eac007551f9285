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
