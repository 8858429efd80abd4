#!/usr/bin/env python3
"""Writes erfcx_table.h: piecewise polynomials for the scaled complementary
error function erfcx(x) = exp(x^2) erfc(x) on [0, 6), used by the pair kernel
to form erfc(alpha r) = exp(-alpha^2 r^2) erfcx(alpha r) from the same
exponential that the damped-shifted-force term needs.

Method: on each of 24 intervals of width 1/4, Chebyshev interpolation of
degree 11 at Chebyshev-Gauss nodes, converted to a polynomial in the
interval variable t = (x - mid) * 8 in [-1, 1] (Horner form).  erfcx is
evaluated in 60-digit decimal arithmetic from the everywhere-positive series
erf(x) = 2/sqrt(pi) exp(-x^2) sum_n 2^n x^(2n+1) / (2n+1)!!  (Abramowitz and
Stegun 7.1.6), so the fit does not depend on any library erfc.  The script
checks the double-precision Horner evaluation against the decimal values on
a dense grid and fails above 3 ulp.

    python capacmd_amd/csrc/gen_erfcx.py      # rewrites erfcx_table.h beside it
"""
from decimal import Decimal as D, getcontext
from fractions import Fraction as F
import os
import struct

getcontext().prec = 60
N_PIECES, WIDTH, DEG = 24, D(1) / 4, 11
SMALL_TERMS = 12


def pi():
    # Machin: pi = 16 atan(1/5) - 4 atan(1/239)
    def atan_inv(n):
        x = D(1) / n
        x2 = x * x
        s, term, k = D(0), x, 0
        while True:
            s += term / (2 * k + 1) * (-1) ** k
            k += 1
            term *= x2
            if term < D(10) ** -70:
                return s
    return 16 * atan_inv(5) - 4 * atan_inv(239)


PI = pi()
SQRTPI = PI.sqrt()


def cos(x):
    x = D(x)
    s, term, k = D(0), D(1), 0
    while abs(term) > D(10) ** -65:
        s += term
        k += 2
        term *= -x * x / ((k - 1) * k)
    return s


def erfcx(x):
    x = D(x)
    if x == 0:
        return D(1)
    s, term, n = D(0), x, 0
    while True:
        s += term
        n += 1
        term = term * 2 * x * x / (2 * n + 1)
        if term < s * D(10) ** -50:
            break
    erf = 2 / SQRTPI * (-(x * x)).exp() * s
    return (x * x).exp() * (1 - erf)


def piece(k):
    a = WIDTH * k
    mid, half = a + WIDTH / 2, WIDTH / 2
    n = DEG + 1
    nodes = [cos(PI * (j + D(1) / 2) / n) for j in range(n)]
    f = [erfcx(mid + half * t) for t in nodes]
    cheb = []
    for m in range(n):
        s = sum(f[j] * cos(PI * m * (j + D(1) / 2) / n) for j in range(n))
        cheb.append(2 * s / n)
    cheb[0] /= 2
    # Chebyshev -> monomial in t: T_0 = 1, T_1 = t, T_{m+1} = 2 t T_m - T_{m-1}
    T = [[D(1)] + [D(0)] * DEG, [D(0), D(1)] + [D(0)] * (DEG - 1)]
    for m in range(2, n):
        nxt = [D(0)] * n
        for i in range(n):
            if i > 0:
                nxt[i] += 2 * T[m - 1][i - 1]
            nxt[i] -= T[m - 2][i]
        T.append(nxt)
    mono = [sum(cheb[m] * T[m][i] for m in range(n)) for i in range(n)]
    return mid, [float(c) for c in mono]


def erf_d(x):
    # Maclaurin series 2/sqrt(pi) sum_n (-1)^n x^(2n+1) / (n! (2n+1)), 60 digits
    x = D(x)
    s, term, n = D(0), x, 0
    while abs(term) > D(10) ** -70:
        s += term / (2 * n + 1)
        n += 1
        term = -term * x * x / n
    return 2 / SQRTPI * s


def erf_small_coef():
    """erf(x) = x p(x^2) on [0, 1/2): p the Maclaurin series cut after
    SMALL_TERMS terms (the first one left out is below 1e-17 of erf there)."""
    c, fact = [], D(1)
    for n in range(SMALL_TERMS):
        if n:
            fact *= n
        c.append(float(2 / SQRTPI * (-1) ** n / (fact * (2 * n + 1))))
    return c


def fma_exact(a, b, c):
    return float(F(a) * F(b) + F(c))


def erf_small_eval(c, x):
    x2 = x * x
    p = c[-1]
    for v in reversed(c[:-1]):
        p = fma_exact(p, x2, v)
    return x * p


def horner(coef, t):
    r = coef[-1]
    for c in reversed(coef[:-1]):
        r = r * t + c                      # the kernel uses fma; the check bounds both
    return r


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    pieces = [piece(k) for k in range(N_PIECES)]
    worst = 0.0
    for k, (mid, coef) in enumerate(pieces):
        for s in range(161):
            x = float(WIDTH * k) + float(WIDTH) * s / 160.0
            if x >= float(WIDTH) * N_PIECES:
                continue
            t = (x - float(mid)) * 8.0
            got = horner(coef, t)
            want = erfcx(D(x))
            err = abs((D(got) - want) / want)
            ulp = D(2) ** (struct.unpack("<q", struct.pack("<d", float(want)))[0] >> 52 & 0x7FF) / D(2) ** (1023 + 52)
            worst = max(worst, float(abs(D(got) - want) / ulp))
    print("largest error: %.2f ulp" % worst)
    if worst > 3.0:
        raise SystemExit("fit not accurate enough")
    small = erf_small_coef()
    worst_small = 0.0
    for s in range(1, 4001):
        x = 0.5 * s / 4001.0
        want = erf_d(D(x))
        ulp = D(2) ** (struct.unpack("<q", struct.pack("<d", float(want)))[0] >> 52 & 0x7FF) / D(2) ** (1023 + 52)
        worst_small = max(worst_small, float(abs(D(erf_small_eval(small, x)) - want) / ulp))
    print("erf series on (0, 1/2): largest error %.2f ulp" % worst_small)
    if worst_small > 2.0:
        raise SystemExit("series not accurate enough")
    lines = [
        "// Generated by gen_erfcx.py (same directory); do not edit.",
        "// erfcx(x) = exp(x^2) erfc(x) on [0, %d): piece k covers [k/4, (k+1)/4)," % (N_PIECES // 4),
        "// erfcx = sum_i c[k][i] t^i with t = (x - (k + 1/2)/4) * 8.  Largest error of",
        "// the double Horner evaluation against 60-digit values: %.2f ulp." % worst,
        "#pragma once",
        "",
        "constexpr int kErfcxPieces = %d;" % N_PIECES,
        "constexpr int kErfcxTerms = %d;" % (DEG + 1),
        "constexpr double kErfcxHi = %d.0;" % (N_PIECES // 4),
        "static __constant__ const double kErfcxCoef[kErfcxPieces][kErfcxTerms] = {",
    ]
    for k, (mid, coef) in enumerate(pieces):
        lines.append("    {" + ", ".join("%.17g" % c for c in coef) + "},")
    lines.append("};")
    lines += [
        "",
        "// erf(x) = x p(x^2) on [0, 1/2): the Maclaurin series of erf,",
        "// 2/sqrt(pi) sum_n (-1)^n x^(2n+1) / (n! (2n+1)), cut after %d terms (Horner" % SMALL_TERMS,
        "// in x^2 with FMA).  Largest error of the double evaluation against",
        "// 60-digit values: %.2f ulp." % worst_small,
        "constexpr int kErfSmallTerms = %d;" % SMALL_TERMS,
        "constexpr double kErfSmall[kErfSmallTerms] = {" + ", ".join("%.17g" % c for c in small) + "};",
    ]
    with open(os.path.join(here, "erfcx_table.h"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
