#!/usr/bin/env python3
"""Writes exp_poly.h: the exponential the pair kernels need (the
damped-shifted-force Gaussian, the erf-damped vdW and Buckingham terms), as

    k = rint(x / ln 2),   f = x - k ln 2  (two FMA steps: ln 2 = L1 + L2),
    exp(x) = 2^k p(f),    |f| <= ln 2 / 2,

(for -700 <= x <= 700: 2^k stays a normal number, no argument checks)

with p the degree-11 Chebyshev interpolant of exp on [-ln 2 / 2, ln 2 / 2]
(Chebyshev-Gauss nodes), converted to monomial form.  Everything is computed
in 60-digit decimal arithmetic (exp from its Taylor series), so the fit does
not depend on any library exp.  Its coefficients are the same for every lane,
so a loop keeps them in registers, and each Horner step is one FMA that
leaves its coefficient's register intact.

The script checks the double-precision evaluation (FMA emulated exactly with
fractions) against the decimal values over [-40, 40] and fails above 2 ulp.

    python capacmd_amd/csrc/gen_exp.py      # rewrites exp_poly.h beside it
"""
from decimal import Decimal as D, getcontext
from fractions import Fraction as F
import math
import os
import random

getcontext().prec = 60
DEG = 11


def exp_d(x):
    x = D(x)
    # exp(x) = exp(x / 2^s)^(2^s) with a small argument for the series
    s = 0
    while abs(x) > D("0.01"):
        x /= 2
        s += 1
    total, term, n = D(0), D(1), 0
    while abs(term) > D(10) ** -70:
        total += term
        n += 1
        term = term * x / n
    for _ in range(s):
        total *= total
    return total


def ln2_d():
    # ln 2 = sum 1 / (k 2^k)
    total, k = D(0), 1
    while True:
        term = D(1) / (k * D(2) ** k)
        if term < D(10) ** -70:
            return total
        total += term
        k += 1


def pi_d():
    def atan_inv(n):
        x = D(1) / n
        s, term, k = D(0), x, 0
        while abs(term) > D(10) ** -70:
            s += term / (2 * k + 1) * (-1) ** k
            k += 1
            term *= x * x
        return s
    return 16 * atan_inv(5) - 4 * atan_inv(239)


def cos_d(x):
    x = D(x)
    s, term, k = D(0), D(1), 0
    while abs(term) > D(10) ** -65:
        s += term
        k += 2
        term *= -x * x / ((k - 1) * k)
    return s


LN2 = ln2_d()
PI = pi_d()
HALF = LN2 / 2


def cheb_monomial():
    """Degree-DEG interpolant of exp on [-HALF, HALF] at Chebyshev-Gauss
    nodes, as monomial coefficients in f (Newton divided differences, then
    expansion, all in decimal)."""
    n = DEG + 1
    xs = [HALF * cos_d(PI * (2 * i + 1) / (2 * n)) for i in range(n)]
    ys = [exp_d(x) for x in xs]
    coef = ys[:]
    for j in range(1, n):
        for i in range(n - 1, j - 1, -1):
            coef[i] = (coef[i] - coef[i - 1]) / (xs[i] - xs[i - j])
    poly = [D(0)] * n                 # Newton form -> monomial
    for i in range(n - 1, -1, -1):
        new = [D(0)] * n
        for k in range(n - 1):
            new[k + 1] += poly[k]
            new[k] -= poly[k] * xs[i]
        new[0] += coef[i]
        poly = new
    return poly


def fma(a, b, c):
    return float(F(a) * F(b) + F(c))


def main():
    L1 = float(LN2)
    L2 = float(LN2 - D(L1))
    inv_ln2 = float(1 / LN2)
    c = [float(v) for v in cheb_monomial()]

    def ev(x):
        k = float(round(x * inv_ln2))          # rint (ties never occur on the check grid)
        f = fma(k, -L1, x)
        f = fma(k, -L2, f)
        p = c[DEG]
        for i in range(DEG - 1, -1, -1):
            p = fma(p, f, c[i])
        return math.ldexp(p, int(k))

    rng = random.Random(7)
    worst = 0.0
    pts = [-40.0 + 80.0 * i / 8000 for i in range(8001)] + [80.0 * rng.random() - 40.0 for _ in range(8000)]
    for x in pts:
        want = exp_d(D(x))
        got = ev(x)
        ulp = math.ulp(float(want))
        worst = max(worst, float(abs(D(got) - want)) / ulp)
    print(f"largest error over [-40, 40]: {worst:.2f} ulp")
    assert worst <= 2.0, worst

    def lit(v):
        return repr(v) if "e" in repr(v) or "." in repr(v) else repr(v) + ".0"

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "exp_poly.h"), "w") as fh:
        fh.write("// Generated by gen_exp.py (same directory); do not edit.\n")
        fh.write("// exp(x) = 2^k p(f), k = rint(x / ln 2), f = x - k ln 2 (ln 2 = kExpL1 + kExpL2),\n")
        fh.write("// p the degree-11 Chebyshev interpolant of exp on [-ln 2 / 2, ln 2 / 2].\n")
        fh.write(f"// Largest error of the double evaluation over [-40, 40] against 60-digit values: {worst:.2f} ulp.\n")
        fh.write("#pragma once\n\n")
        fh.write(f"constexpr double kExpInvLn2 = {lit(inv_ln2)};\n")
        fh.write(f"constexpr double kExpL1 = {lit(L1)};\n")
        fh.write(f"constexpr double kExpL2 = {lit(L2)};\n")
        fh.write(f"constexpr int kExpDeg = {DEG};\n")
        fh.write("constexpr double kExpCoef[kExpDeg + 1] = {\n")
        for v in c:
            fh.write(f"    {lit(v)},\n")
        fh.write("};\n\n")
        fh.write("// a * b + c as one v_fma_f64 whose addend register is left intact: in a\n")
        fh.write("// loop the compiler keeps each coefficient in a register, and its own\n")
        fh.write("// choice, v_fmac_f64 (which overwrites the addend), costs a copy per step\n")
        fh.write("__device__ __forceinline__ double fma_keep_addend(double a, double b, double c) {\n")
        fh.write("    double d;\n")
        fh.write('    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));\n')
        fh.write("    return d;\n")
        fh.write("}\n\n")
        fh.write("// exp(x) for -700 <= x <= 700 (no checks for other arguments).  Keep: each\n")
        fh.write("// Horner step through fma_keep_addend (for loops with registers to spare);\n")
        fh.write("// otherwise plain fma, the compiler's choice of registers and instructions.\n")
        fh.write("template <bool Keep>\n")
        fh.write("__device__ __forceinline__ double exp_poly(double x) {\n")
        fh.write("    const double k = rint(x * kExpInvLn2);\n")
        fh.write("    double f = fma(k, -kExpL1, x);\n")
        fh.write("    f = fma(k, -kExpL2, f);\n")
        fh.write("    double p = kExpCoef[kExpDeg];\n")
        fh.write("#pragma unroll\n")
        fh.write("    for (int i = kExpDeg - 1; i >= 0; --i) p = Keep ? fma_keep_addend(p, f, kExpCoef[i]) : fma(p, f, kExpCoef[i]);\n")
        fh.write("    return ldexp(p, (int)k);\n")
        fh.write("}\n")


if __name__ == "__main__":
    main()
