// Generated by gen_exp.py (same directory); do not edit.
// exp(x) = 2^k p(f), k = rint(x / ln 2), f = x - k ln 2 (ln 2 = kExpL1 + kExpL2),
// p the degree-11 Chebyshev interpolant of exp on [-ln 2 / 2, ln 2 / 2].
// Largest error of the double evaluation over [-40, 40] against 60-digit values: 0.86 ulp.
#pragma once

constexpr double kExpInvLn2 = 1.4426950408889634;
constexpr double kExpL1 = 0.6931471805599453;
constexpr double kExpL2 = 2.3190468138462996e-17;
constexpr int kExpDeg = 11;
constexpr double kExpCoef[kExpDeg + 1] = {
    1.0,
    1.0,
    0.5000000000000019,
    0.1666666666666668,
    0.0416666666664881,
    0.008333333333319601,
    0.0013888888952314775,
    0.00019841269890047113,
    2.4801485482328494e-05,
    2.755724091857897e-06,
    2.763263963904103e-07,
    2.5110037605963777e-08,
};

// a * b + c as one v_fma_f64 whose addend register is left intact: in a
// loop the compiler keeps each coefficient in a register, and its own
// choice, v_fmac_f64 (which overwrites the addend), costs a copy per step
__device__ __forceinline__ double fma_keep_addend(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// exp(x) for -700 <= x <= 700 (no checks for other arguments).  Keep: each
// Horner step through fma_keep_addend (for loops with registers to spare);
// otherwise plain fma, the compiler's choice of registers and instructions.
template <bool Keep>
__device__ __forceinline__ double exp_poly(double x) {
    const double k = rint(x * kExpInvLn2);
    double f = fma(k, -kExpL1, x);
    f = fma(k, -kExpL2, f);
    double p = kExpCoef[kExpDeg];
#pragma unroll
    for (int i = kExpDeg - 1; i >= 0; --i) p = Keep ? fma_keep_addend(p, f, kExpCoef[i]) : fma(p, f, kExpCoef[i]);
    return ldexp(p, (int)k);
}
