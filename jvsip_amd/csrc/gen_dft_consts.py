"""Generate hip/dft_consts.h: cos/sin of 2*pi*m/P for the odd prime butterflies.

The constants are written as 20-significant-digit decimal literals computed with
Python's arbitrary-precision Decimal arithmetic, so both the float and the double
kernels round them correctly at compile time.  Run from this directory:
    python gen_dft_consts.py > hip/dft_consts.h
"""
from decimal import Decimal, getcontext

getcontext().prec = 50
PRIMES = (3, 5, 7, 11, 13, 17, 19, 23, 29, 31)
# composite radices built from two smaller DFTs; their internal twiddles
COMPOSITES = (6, 9, 10, 12, 14, 15)


def _pi():
    # Machin's formula: pi = 16 atan(1/5) - 4 atan(1/239)
    def atan_inv(x):
        x = Decimal(x)
        total, term, k, sign = Decimal(0), 1 / x, 1, 1
        x2 = x * x
        while term > Decimal(10) ** -48:
            total += sign * term / k
            term /= x2
            k += 2
            sign = -sign
        return total
    return 16 * atan_inv(5) - 4 * atan_inv(239)


def _cos_sin(theta):
    # Taylor series; theta is in [0, 2*pi) and the precision is ample for it.
    c, s = Decimal(0), Decimal(0)
    term, k = Decimal(1), 0
    while abs(term) > Decimal(10) ** -48 or k < 4:
        if k % 4 == 0:
            c += term
        elif k % 4 == 1:
            s += term
        elif k % 4 == 2:
            c -= term
        else:
            s -= term
        k += 1
        term = term * theta / k
    return c, s


def main():
    pi = _pi()
    out = ["/* Generated by gen_dft_consts.py -- do not edit. */",
           "#pragma once",
           "template <int P> struct PrimeTrig;"]
    for p in PRIMES:
        cs, ss = [], []
        for m in range(p):
            c, s = _cos_sin(2 * pi * m / p)
            cs.append("%.20e" % c)
            ss.append("%.20e" % s)
        out.append("template <> struct PrimeTrig<%d> {" % p)
        out.append("  /* cos(2 pi m / %d), sin(2 pi m / %d), m = 0..%d */" % (p, p, p - 1))
        out.append("  static constexpr double c[%d] = {%s};" % (p, ", ".join(cs)))
        out.append("  static constexpr double s[%d] = {%s};" % (p, ", ".join(ss)))
        out.append("};")
    out.append("template <int R> struct CompTrig;")
    for r in COMPOSITES:
        cs, ss = [], []
        for m in range(r):
            c, s_ = _cos_sin(2 * pi * m / r)
            cs.append("%.20e" % c)
            ss.append("%.20e" % s_)
        out.append("template <> struct CompTrig<%d> {" % r)
        out.append("  /* cos(2 pi m / %d), sin(2 pi m / %d), m = 0..%d */" % (r, r, r - 1))
        out.append("  static constexpr double c[%d] = {%s};" % (r, ", ".join(cs)))
        out.append("  static constexpr double s[%d] = {%s};" % (r, ", ".join(ss)))
        out.append("};")
    print("\n".join(out))


if __name__ == "__main__":
    main()
