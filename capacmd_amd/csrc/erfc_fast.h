// erfc(x) from exp(-x^2), which the damped-shifted-force terms need anyway:
// erfc = exp(-x^2) erfcx(x), erfcx from the piecewise polynomials of
// erfcx_table.h on [0, 6) (within 1 ulp of 60-digit values), the library
// erfc elsewhere.  With exp(-alpha^2 r^2) in place of exp(-(alpha r)^2) the
// exponent differs by rounding only (a relative 1e-15 in erfc).  coef is the
// table staged in LDS by load_erfcx_table.
#pragma once
#include "erfcx_table.h"

__device__ __forceinline__ double erfc_from_exp(double x, double ex, const double* __restrict__ coef) {
    if (!(x >= 0.0 && x < kErfcxHi)) return erfc(x);
    const int k = (int)(x * 4.0);
    const double t = (x - (k + 0.5) * 0.25) * 8.0;
    const double* c = coef + k * kErfcxTerms;
    double p = c[kErfcxTerms - 1];
#pragma unroll
    for (int i = kErfcxTerms - 2; i >= 0; --i) p = fma(p, t, c[i]);
    return ex * p;
}

// erf(x) for 0 <= x < 1/2 from its Maclaurin series (kErfSmall)
__device__ __forceinline__ double erf_small(double x) {
    const double x2 = x * x;
    double p = kErfSmall[kErfSmallTerms - 1];
#pragma unroll
    for (int i = kErfSmallTerms - 2; i >= 0; --i) p = fma(p, x2, kErfSmall[i]);
    return x * p;
}

// erf(x) = 1 - exp(-x^2) erfcx(x) for 0.5 <= x < 6, from the exponential
// the caller already holds (as erfc_from_exp); there erfc <= 0.48, so the
// difference loses nothing: within 2 ulp of the library erf over a 2e5-point
// scan of [0.5, 6).  Beyond 6
// erfc < 2.2e-17 and the result is 1.0, as the library's.  Below 0.5 (where
// 1 - erfc would cancel) the Maclaurin series; for negative or NaN x the
// library erf.
__device__ __forceinline__ double erf_from_exp(double x, double ex, const double* __restrict__ coef) {
    if (!(x >= 0.5)) return x >= 0.0 ? erf_small(x) : erf(x);
    if (!(x < kErfcxHi)) return 1.0;
    const int k = (int)(x * 4.0);
    const double t = (x - (k + 0.5) * 0.25) * 8.0;
    const double* c = coef + k * kErfcxTerms;
    double p = c[kErfcxTerms - 1];
#pragma unroll
    for (int i = kErfcxTerms - 2; i >= 0; --i) p = fma(p, t, c[i]);
    return 1.0 - ex * p;
}

// erf(x) for x >= 0 (not NaN) without the library fallback: the Maclaurin
// series below 1/2, 1 - exp(-x^2) erfcx(x) up to 6, 1 beyond
__device__ __forceinline__ double erf_from_exp_pos(double x, double ex, const double* __restrict__ coef) {
    if (x < 0.5) return erf_small(x);
    if (!(x < kErfcxHi)) return 1.0;
    const int k = (int)(x * 4.0);
    const double t = (x - (k + 0.5) * 0.25) * 8.0;
    const double* c = coef + k * kErfcxTerms;
    double p = c[kErfcxTerms - 1];
#pragma unroll
    for (int i = kErfcxTerms - 2; i >= 0; --i) p = fma(p, t, c[i]);
    return 1.0 - ex * p;
}

// the whole workgroup copies the table into s (kErfcxPieces * kErfcxTerms
// doubles of LDS); the caller synchronises before use
__device__ __forceinline__ void load_erfcx_table(double* s) {
    for (int k = threadIdx.x; k < kErfcxPieces * kErfcxTerms; k += blockDim.x)
        s[k] = kErfcxCoef[k / kErfcxTerms][k % kErfcxTerms];
}
