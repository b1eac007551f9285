"""CPU model of the three-pass split (no device needed).

The library splits long transforms, MN = n1 n2 n3, into three launches whose
layouts are fixed by jvsip_amd/csrc/host/va_fftm.c (va_descs3): the vector
numbering, element strides, scratch layouts and post-twiddles of each pass.
This restates exactly those layouts in numpy and checks the composition
against numpy.fft, with planted mistakes (a twiddle taken modulo the wrong
count, a transposed scratch layout) that must fail.
"""
import numpy as np
import pytest


def three_pass(x, n1, n2, n3, ptw2_mod=None, t2_swap=False):
    mN, MN = x.shape
    n23 = n2 * n3
    t1 = np.zeros(mN * MN, complex)
    t2 = np.zeros(mN * MN, complex)
    out = np.zeros(mN * MN, complex)
    xf = x.ravel()
    w = lambda a, n: np.exp(-2j * np.pi * a / n)
    for v in range(mN * n23):                         # pass 1
        row, m = divmod(v, n23)
        idx = row * MN + m + n23 * np.arange(n1)
        t1[idx] = np.fft.fft(xf[idx]) * w((v % n23) * np.arange(n1), MN)
    mod = ptw2_mod or n3
    for v in range(mN * n1 * n3):                     # pass 2
        ino = (v % n3) + (v // n3) * n23
        if t2_swap:
            outo, es = (v % n3) + (v // n3) * n23, n3
        else:
            outo, es = (v % (n1 * n3)) + (v // (n1 * n3)) * MN, n1 * n3
        y = np.fft.fft(t1[ino + n3 * np.arange(n2)]) * w((v % mod) * np.arange(n2), n23)
        t2[outo + es * np.arange(n2)] = y
    for v in range(mN * n1 * n2):                     # pass 3
        outo = (v % (n1 * n2)) + (v // (n1 * n2)) * MN
        out[outo + n1 * n2 * np.arange(n3)] = np.fft.fft(t2[v * n3 + np.arange(n3)])
    return out.reshape(mN, MN)


@pytest.mark.parametrize("n1,n2,n3,m", [(2, 3, 4, 2), (4, 4, 8, 3), (5, 3, 7, 1), (8, 8, 16, 2)])
def test_three_pass_layouts_compose_to_the_dft(n1, n2, n3, m):
    rng = np.random.default_rng(n1 * 100 + n3)
    x = rng.standard_normal((m, n1 * n2 * n3)) + 1j * rng.standard_normal((m, n1 * n2 * n3))
    err = np.abs(three_pass(x, n1, n2, n3) - np.fft.fft(x, axis=1)).max()
    assert err < 1e-12


def test_three_pass_planted_mistakes_fail():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((2, 4 * 3 * 5)) + 1j * rng.standard_normal((2, 60))
    want = np.fft.fft(x, axis=1)
    assert np.abs(three_pass(x, 4, 3, 5, ptw2_mod=4 * 5) - want).max() > 1e-3
    assert np.abs(three_pass(x, 4, 3, 5, t2_swap=True) - want).max() > 1e-3
