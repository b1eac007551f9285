"""CPU checks of the k-omega beamformer pieces (no device needed).

The oracle's komega() (oracle/pyoracle.py) is pinned against the
reference's own C, compiled from its sources into oracle/_ref, running
kw.c's komega() call for call (jvsip_amd/komega.py drives any library that
exports the API), and against a plane-wave closed form; the Hanning taper
against the reference's vsip_vcreate_hanning_f/_d bit for bit.
"""
import numpy as np
import pytest

from jvsip_amd import komega as kw
from oracle import pyoracle as po


@pytest.mark.parametrize("n", [1, 2, 7, 128, 1000, 1024])
def test_hanning_matches_reference_bit_for_bit(ref, n):
    for t, dt in (("f", np.float32), ("d", np.float64)):
        v = getattr(ref, "vcreate_hanning_" + t)(n, 0)
        got = np.empty(n, dt)
        getattr(ref, "vcopyto_user_" + t)(v, got.ctypes.data)
        getattr(ref, "valldestroy_" + t)(v)
        np.testing.assert_array_equal(got, po.hanning(n, t))


def test_hanning_closed_form():
    w = po.hanning(9, "d")
    np.testing.assert_allclose(w, 0.5 * (1 - np.cos(2 * np.pi * np.arange(1, 10) / 10)), rtol=0, atol=1e-13)
    np.testing.assert_allclose(w, w[::-1], rtol=0, atol=1e-13)       # symmetric
    assert abs(w[4] - 1.0) < 1e-13                                     # centre of an odd window


@pytest.mark.parametrize("m,n,navg", [(8, 16, 1), (16, 24, 2), (128, 1024, 3)])
def test_oracle_komega_matches_reference_kw(ref, m, n, navg):
    k = kw.KOmega(ref, m, n, navg)
    k.zero()
    gram = np.zeros((m, n // 2 + 1))
    ts, arr = po.hanning(n, "f"), po.hanning(m, "f")
    for it in range(navg):
        x = kw.plane_waves(m, n, [(n // 8, 1, 1.0), (n // 3, m // 2 - 1, 0.5)], seed=10 + it, noise=0.3)
        dm = kw.data_matrix(ref, x)
        k.step(dm)
        tapered = kw.read_matrix(ref, dm, (m, n))
        np.testing.assert_allclose(tapered, x * ts[None, :] * arr[:, None], rtol=2e-7, atol=0)
        ref.malldestroy_f(dm)
        gram = po.komega(x, ts, arr, gram, 1.0 / navg)
    got = k.gram()
    k.destroy()
    err = po.rel_l2(got, gram)
    assert err < 1e-5, err


def test_oracle_komega_plane_wave_peak():
    """A plane wave at time bin f and sensor bin k: the gram peaks at
    frequency f and direction bin -k mod M (kw's column transform has the
    inverse sign)."""
    m, n = 32, 256
    for f, k in ((20, 3), (77, 10), (5, 0)):
        x = kw.plane_waves(m, n, [(f, k, 1.0)])
        g = po.komega(x, po.hanning(n, "d"), po.hanning(m, "d"), np.zeros((m, n // 2 + 1)), 1.0)
        i, j = np.unravel_index(np.argmax(g), g.shape)
        assert (j, i) == (f, (-k) % m), (f, k, i, j)
        # closed form of the peak: cos = (e^{i phi} + e^{-i phi}) / 2 gives
        # (Ts As + T2 A2) / 2 with Ts, As the taper sums and T2, A2 their
        # transforms at twice the tone's bins
        ts, arr = po.hanning(n, "d"), po.hanning(m, "d")
        t2 = np.sum(ts * np.exp(-4j * np.pi * f * np.arange(n) / n))
        a2 = np.sum(arr * np.exp(-4j * np.pi * k * np.arange(m) / m))
        want = abs(0.5 * (ts.sum() * arr.sum() + t2 * a2)) ** 2
        assert abs(g[i, j] - want) <= 1e-6 * want, (g[i, j], want)       # data rounded to float32
