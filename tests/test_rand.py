"""CPU checks of the VSIPL random number generators (no device needed).

vsip_randcreate and the scalar draws (randu/randn/crandu/crandn, _f and _d)
run on the host; they must give the reference's values bit for bit (its own
C, compiled from its sources into oracle/_ref) for both generators, several
(seed, numseqs, id), and across the portable generator's marker skip, which
the tests reach by placing X1 just before X2 in both states (the state
struct has the reference's field order, vsip_randobject.h).
"""
import ctypes

import numpy as np
import pytest

from jvsip_amd.vsip import VSIP_MEM_NONE

A1 = 69069


class RandState(ctypes.Structure):
    _fields_ = [("a", ctypes.c_uint), ("c", ctypes.c_uint), ("a1", ctypes.c_uint), ("c1", ctypes.c_uint),
                ("X", ctypes.c_uint), ("X1", ctypes.c_uint), ("X2", ctypes.c_uint), ("type", ctypes.c_int)]


def state(h):
    return ctypes.cast(h, ctypes.POINTER(RandState)).contents


def lcg_back(a, c, x, k):
    """x after k steps backwards of x <- a x + c (mod 2^32)"""
    ai = pow(a, -1, 1 << 32)
    for _ in range(k):
        x = (ai * (x - c)) % (1 << 32)
    return x


@pytest.fixture(scope="module")
def lib():
    from jvsip_amd.vsip import AmdLib, LIB_PATH
    return AmdLib(LIB_PATH)


def _val(v):
    return (v.r, v.i) if hasattr(v, "r") else v


CREATE = [(0, 1, 1), (17, 1, 1), (12345, 4, 3), (99, 7, 7), (2 ** 31 + 5, 100, 100), (7, 3, 1), (5, 1000003, 77)]


@pytest.mark.parametrize("seed,numseqs,ident", CREATE)
@pytest.mark.parametrize("portable", [0, 1])
def test_scalar_draws_match_reference(lib, ref, seed, numseqs, ident, portable):
    a = lib.randcreate(seed, numseqs, ident, 1 - portable)
    b = ref.randcreate(seed, numseqs, ident, 1 - portable)
    assert a and b
    sa, sb = state(a), state(b)
    # the reference leaves a1, c1, X1, X2 unset for the nonportable generator
    for f in ("a", "c", "X", "type") + (("a1", "c1", "X1", "X2") if portable else ()):
        assert getattr(sa, f) == getattr(sb, f), f
    for name in ("randu_f", "randn_f", "crandu_f", "crandn_f", "randu_d", "randn_d", "crandu_d", "crandn_d"):
        for _ in range(50):
            assert _val(getattr(lib, name)(a)) == _val(getattr(ref, name)(b)), name
    assert sa.X == sb.X and (not portable or (sa.X1, sa.X2) == (sb.X1, sb.X2))
    assert lib.randdestroy(a) == 0
    ref.randdestroy(b)


def test_marker_skip_matches_reference(lib, ref):
    """X1 meets X2 within the next draws: both step past it the same way."""
    for back in (1, 2, 5, 11):
        a, b = lib.randcreate(5, 1, 1, 0), ref.randcreate(5, 1, 1, 0)
        for h in (a, b):
            s = state(h)
            s.X1 = lcg_back(A1, s.c1, s.X2, back)
        for _ in range(8):
            assert _val(lib.crandn_f(a)) == _val(ref.crandn_f(b))
            assert _val(lib.randu_d(a)) == _val(ref.randu_d(b))
        sa, sb = state(a), state(b)
        assert (sa.X, sa.X1, sa.X2) == (sb.X, sb.X1, sb.X2) and sa.X2 == 2
        lib.randdestroy(a)
        ref.randdestroy(b)


def test_create_refuses_undefined_arguments(lib):
    assert not lib.randcreate(1, 0, 1, 0)        # numseqs 0
    assert not lib.randcreate(1, 4, 0, 0)        # portable id 0
    assert not lib.randcreate(1, 200, 101, 0)    # past the 100 primes
    h = lib.randcreate(1, 4, 0, 1)               # nonportable id 0 is the seed itself
    assert h and state(h).X == 1
    lib.randdestroy(h)


def test_uniform_moments(lib):
    """Large-sample sanity of the values (mean 1/2, var 1/12; normal 0, 1)."""
    h = lib.randcreate(3, 1, 1, 0)
    u = np.array([lib.randu_f(h) for _ in range(20000)])
    n = np.array([lib.randn_d(h) for _ in range(20000)])
    lib.randdestroy(h)
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    assert abs(n.mean()) < 0.03 and abs(n.var() - 1) < 0.05
