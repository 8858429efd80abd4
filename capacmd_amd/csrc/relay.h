// Relay-matrix blocks of one metal pair, shared by the assembly (cpim.hip),
// the SpMV (solver.hip) and the dense expansion used by the parity tests.
//
// Every ordered metal pair (i, j), i != j, owns one 4x4 block
//     [ A_ab (3x3)   M_a = c3 r_a ]        rows: dipole p_i (a), charge q_i
//     [ -c3 r_b      C            ]        cols: dipole p_j (b), charge q_j
// with r = (r_j - r_i) NM2BOHR (cpff.c:259-263) and the Gaussian-smeared
// T-tensor coefficients (cpff.c:266-287)
//   c1 = [erf(r/R_pp) - 2/sqrt(pi) (r/R_pp) e^{-r^2/R_pp^2}] / r^5
//   c2 = 4/sqrt(pi) R_pp^-3 e^{-r^2/R_pp^2} / r^2
//   c3 = [erf(r/R_pq) - 2/sqrt(pi) (r/R_pq) e^{-r^2/R_pq^2}] / r^3
//   cq = -erf(r/R_qq) / r
// so that
//   A_ab = -(c1 (3 r_a r_b - delta_ab r^2) - c2 r_a r_b)    (cpff.c:326-328)
//        = g_a r_a r_b + g_b delta_ab,  g_a = c2 - 3 c1,  g_b = c1 r^2
//   M_a  = c3 r_a (cpff.c:353),  M^T_b = -c3 r_b (cpff.c:386),  C = cq (cpff.c:428).
// The diagonal blocks are diag(1/alpha) and -1/c (cpff.c:305,408).
#pragma once

#include "capac_internal.h"
#include "erfc_fast.h"
#include "exp_poly.h"

struct TCoef {
    double c1, c2, c3, cq;
};

__device__ __forceinline__ TCoef t_coef(const BoxConst& bc, double r2) {
    TCoef t;
    double rij = sqrt(r2);
    double inv_r = 1.0 / rij;
    double inv_r2 = inv_r * inv_r;
    double inv_r3 = inv_r2 * inv_r;
    double e_pp = exp(-(r2 * bc.inv_R_pp * bc.inv_R_pp));
    double inv_Rpp3 = bc.inv_R_pp * bc.inv_R_pp * bc.inv_R_pp;
    t.c1 = inv_r3 * inv_r2 * (erf(rij * bc.inv_R_pp) - 2.0 * K_INV_SQRT_PI * bc.inv_R_pp * rij * e_pp);
    t.c2 = 4.0 * K_INV_SQRT_PI * inv_Rpp3 * inv_r2 * e_pp;
    double e_pq = exp(-(r2 * bc.inv_R_pq * bc.inv_R_pq));
    t.c3 = inv_r3 * (erf(rij * bc.inv_R_pq) - 2.0 * K_INV_SQRT_PI * bc.inv_R_pq * rij * e_pq);
    t.cq = -erf(rij * bc.inv_R_qq) * inv_r;
    return t;
}

// t_coef with each exponential from exp_poly, each erf from the erfcx table
// (erf_from_exp: the Gaussian the coefficient needs anyway; the library erf
// below 0.5) and 1/r from rsqrt_pos: the same expressions, within a few ulp
// of t_coef at a third of its instructions.  erfcx is the table staged in
// LDS (load_erfcx_table).
__device__ __forceinline__ TCoef t_coef_fast(const BoxConst& bc, double r2, const double* __restrict__ erfcx) {
    TCoef t;
    const double inv_r = rsqrt_pos(r2);
    const double rij = r2 * inv_r;
    const double inv_r2 = inv_r * inv_r;
    const double inv_r3 = inv_r2 * inv_r;
    const double e_pp = exp_poly<false>(-(r2 * bc.inv_R_pp * bc.inv_R_pp));
    const double inv_Rpp3 = bc.inv_R_pp * bc.inv_R_pp * bc.inv_R_pp;
    t.c1 = inv_r3 * inv_r2 *
           (erf_from_exp(rij * bc.inv_R_pp, e_pp, erfcx) - 2.0 * K_INV_SQRT_PI * bc.inv_R_pp * rij * e_pp);
    t.c2 = 4.0 * K_INV_SQRT_PI * inv_Rpp3 * inv_r2 * e_pp;
    const double e_pq = exp_poly<false>(-(r2 * bc.inv_R_pq * bc.inv_R_pq));
    t.c3 = inv_r3 * (erf_from_exp(rij * bc.inv_R_pq, e_pq, erfcx) - 2.0 * K_INV_SQRT_PI * bc.inv_R_pq * rij * e_pq);
    const double e_qq = exp_poly<false>(-(r2 * bc.inv_R_qq * bc.inv_R_qq));
    t.cq = -erf_from_exp(rij * bc.inv_R_qq, e_qq, erfcx) * inv_r;
    return t;
}

// separation r_j - r_i of two metal atoms in bohr, the reference's pbc_12
// on the nm difference, then the unit change
__device__ __forceinline__ void metal_sep(const double* __restrict__ r, const BoxConst& bc, int i, int j,
                                          double& dx, double& dy, double& dz) {
    const int N = bc.N;
    dx = r[j] - r[i];
    dy = r[N + j] - r[N + i];
    dz = r[2 * N + j] - r[2 * N + i];
    pbc12(bc, dx, dy, dz);
    dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
}

// FP64 block for the separation (dx, dy, dz) in nm, i != j, the reference's
// expressions term by term
__device__ __forceinline__ void relay_block64_sep(double dx, double dy, double dz, const BoxConst& bc,
                                                  double blk[16]) {
    pbc12(bc, dx, dy, dz);
    dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
    double r2 = dist2_exact(dx, dy, dz);
    TCoef t = t_coef(bc, r2);
    double rv[3] = {dx, dy, dz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
            blk[a * 4 + bb] = -(t.c1 * (3.0 * rv[a] * rv[bb] - (a == bb ? r2 : 0.0)) -
                                t.c2 * rv[a] * rv[bb]);
        blk[a * 4 + 3] = t.c3 * rv[a];
        blk[12 + a] = -t.c3 * rv[a];
    }
    blk[15] = t.cq;
}

// FP64 block (i, j), i != j
__device__ __forceinline__ void relay_block64(const double* __restrict__ r, const BoxConst& bc, int i,
                                              int j, double blk[16]) {
    const int N = bc.N;
    relay_block64_sep(r[j] - r[i], r[N + j] - r[N + i], r[2 * N + j] - r[2 * N + i], bc, blk);
}

// The FP32 relay stores, per ordered pair, the four scalars of |r| that
// the block is made of: g_a, g_b, c3, cq (FP32 T-tensor evaluation), plane k
// of pair (i, j) at relay32[(i * kRelay32Planes + k) * ld + j].  The
// separation r itself is formed again from the FP64 positions wherever a
// block is applied, so 16 bytes per pair carry the whole block
// (kRelay32Planes, capac_internal.h).

// FP32 T-tensor coefficients: the expressions of t_coef in single precision
struct TCoef32 {
    float c1, c2, c3, cq;
};

__device__ __forceinline__ TCoef32 t_coef32(float r2, float iRpp, float iRpq, float iRqq) {
    const float isp = (float)K_INV_SQRT_PI;
    TCoef32 t;
    float rij = sqrtf(r2);
    float inv_r = 1.0f / rij;
    float inv_r2 = inv_r * inv_r;
    float inv_r3 = inv_r2 * inv_r;
    float e_pp = expf(-(r2 * iRpp * iRpp));
    t.c1 = inv_r3 * inv_r2 * (erff(rij * iRpp) - 2.0f * isp * iRpp * rij * e_pp);
    t.c2 = 4.0f * isp * (iRpp * iRpp * iRpp) * inv_r2 * e_pp;
    float e_pq = expf(-(r2 * iRpq * iRpq));
    t.c3 = inv_r3 * (erff(rij * iRpq) - 2.0f * isp * iRpq * rij * e_pq);
    t.cq = -erff(rij * iRqq) * inv_r;
    return t;
}

// block (i, j) applied to (p_j, q_j) from the four coefficients and the
// separation (in FP64 arithmetic):
//   dipole rows (g_a (r.p) + c3 q) r + g_b p,   charge row cq q - c3 (r.p)
__device__ __forceinline__ void block32_apply(double rx, double ry, double rz, double ga, double gb, double c3,
                                              double cq, double x0, double x1, double x2, double x3,
                                              double o[4]) {
    double rp = rx * x0 + ry * x1 + rz * x2;
    double gr = ga * rp + c3 * x3;
    o[0] += gr * rx + gb * x0;
    o[1] += gr * ry + gb * x1;
    o[2] += gr * rz + gb * x2;
    o[3] += cq * x3 - c3 * rp;
}
