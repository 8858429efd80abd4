// Gradients of a force- and energy-matching loss with respect to the CPIM and
// Sutton-Chen parameters, for fitting.  No counterpart in the reference.
//
// Parameters theta = (alpha, c, R_q, R_p, eps_SC, c_SC): polarizability
// (bohr^3) and capacitance (bohr) of the metal atoms, the two Gaussian widths
// (bohr; free parameters here, where the reference derives them from alpha
// and c, mpi-main.c:318-323) and the Sutton-Chen energy scale and embedding
// constant.  One frame's loss is
//   L = sum_{metal i} |F_i - F_ref,i|^2 + w_E (E - E_ref)^2
// with F the forces and E = sum pot[1..14] of one force call, the induced
// charges and dipoles x solving A(theta) x = b.
//
// dL/dtheta by the adjoint method.  A is symmetric, so lambda solves
// A lambda = dL/dx with the same BiCGSTAB kernels, and
//   dL/dtheta = dL/dtheta|_x - lambda . (dA/dtheta) x.
// dL/dx needs G . dF/dx with G_i = 2 (F_i - F_ref,i).  The CPIM forces are
// minus the position gradient of the CPIM energy at fixed x (the reference's
// Hellmann-Feynman forces, cpff.c:526-950), so G . dF/dx = -D_G (dE/dx), the
// derivative of dE/dx when the metal atoms move along G.  That and every
// dtheta term are forward-mode derivatives, evaluated here by re-running the
// energy, force and relay expressions on dual numbers (value, derivative):
// T-tensors are recomputed, never stored.
#include "capac_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

namespace {

constexpr double kTwoOverSqrtPi = 1.1283791670955126;   // d erf(x)/dx = 2/sqrt(pi) e^{-x^2}

struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual dl(double v, double d = 0.0) { return Dual{v, d}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
    double q = a / b.v;
    return {q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
__device__ __forceinline__ Dual dsqrt(Dual a) {
    double s = sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
__device__ __forceinline__ Dual dexp(Dual a) {
    double e = exp(a.v);
    return {e, a.d * e};
}
__device__ __forceinline__ Dual derf(Dual a) { return {erf(a.v), a.d * kTwoOverSqrtPi * exp(-a.v * a.v)}; }
__device__ __forceinline__ Dual derfc(Dual a) { return {erfc(a.v), -a.d * kTwoOverSqrtPi * exp(-a.v * a.v)}; }

__device__ __forceinline__ Dual wave_sum_d(Dual a) { return {wave_sum(a.v), wave_sum(a.d)}; }

// dE/dx (value) and its derivative along G (metal atoms displaced by G, the
// rest fixed), E = pot[9..14] of the CPIM force call (cpff.c:581-948).  One
// wavefront per metal atom i; rows [p_i (3) | q_i] of the vec_pq layout.
__global__ void __launch_bounds__(256)
k_gradx(const double* __restrict__ r, const double* __restrict__ charge, BoxConst bc,
        const int* __restrict__ n_front, const int* __restrict__ nbr, int stride,
        const double* __restrict__ x, const double* __restrict__ G, double* __restrict__ gE,
        double* __restrict__ dgE) {
    const int N = bc.N, Nm = bc.Nm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wave;
    if (i >= Nm) return;                               // no barrier in this kernel
    const double Gi[3] = {G[i], G[Nm + i], G[2 * Nm + i]};
    const double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
    Dual gq = dl(0.0), gp[3] = {dl(0.0), dl(0.0), dl(0.0)};
    // metal - water: pot[10] (charge) and pot[11] (dipole), r = r_metal - r_atom in nm
    const int nf = n_front[i];
    const int* row = nbr + (size_t)i * stride;
    for (int k = lane; k < nf; k += 64) {
        const int a = row[k];
        const double qa = charge[a];
        if (qa == 0.0) continue;
        double dx = xi - r[a], dy = yi - r[N + a], dz = zi - r[2 * N + a];
        pbc12(bc, dx, dy, dz);
        if (!(dist2_exact(dx, dy, dz) < bc.rc2)) continue;
        const Dual rv[3] = {dl(dx, Gi[0]), dl(dy, Gi[1]), dl(dz, Gi[2])};
        const Dual r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        const Dual rij = dsqrt(r2);
        Dual a0;
        if (bc.use_coulomb == 1) {
            const Dual erfc_a = derfc(bc.alpha * rij);
            const Dual e2 = dexp(-(bc.alpha * bc.alpha) * r2);
            gq += K_FQQ * qa * (erfc_a / rij - bc.erfc_arCut + bc.wolfConst * (rij - bc.rc));
            a0 = (erfc_a / r2 + bc.a2_sqrtPI * e2 / rij - bc.wolfConst) / rij;
        } else {
            gq += K_FQQ * qa / rij;
            a0 = 1.0 / (r2 * rij);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) gp[c] -= (K_FQQ * qa / K_NM2BOHR) * (a0 * rv[c]);
    }
    // metal - metal: pot[12..14] pair terms, r = r_j - r_i in bohr
    const double iRqq = bc.inv_R_qq, iRpq = bc.inv_R_pq, iRpp = bc.inv_R_pp;
    const double iRpp3 = iRpp * iRpp * iRpp;
    for (int j = lane; j < Nm; j += 64) {
        if (j == i) continue;
        double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
        pbc12(bc, dx, dy, dz);
        const Dual rv[3] = {dl(dx * K_NM2BOHR, (G[j] - Gi[0]) * K_NM2BOHR),
                            dl(dy * K_NM2BOHR, (G[Nm + j] - Gi[1]) * K_NM2BOHR),
                            dl(dz * K_NM2BOHR, (G[2 * Nm + j] - Gi[2]) * K_NM2BOHR)};
        const Dual r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        const Dual rij = dsqrt(r2);
        const Dual inv_r = 1.0 / rij;
        const Dual inv_r3 = inv_r * inv_r * inv_r;
        const Dual inv_r5 = inv_r3 * inv_r * inv_r;
        const double qj = x[3 * Nm + j];
        const double pj[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
        // pot[12]: 0.5 au1 q_i q_j erf(r/R_qq)/r over both orders -> au1 q_j erf/r
        gq += K_FQQ_AU1 * qj * (derf(rij * iRqq) * inv_r);
        // pot[13]: 0.5 au1 Apq0 (q_j p_i.r - q_i p_j.r) over both orders
        const Dual e_pq = dexp(-(r2 * (iRpq * iRpq)));
        const Dual Apq0 = inv_r3 * (derf(rij * iRpq) - (2.0 * K_INV_SQRT_PI * iRpq) * rij * e_pq);
        const Dual pjr = pj[0] * rv[0] + pj[1] * rv[1] + pj[2] * rv[2];
        gq -= K_FQQ_AU1 * (Apq0 * pjr);
#pragma unroll
        for (int c = 0; c < 3; ++c) gp[c] += (K_FQQ_AU1 * qj) * (Apq0 * rv[c]);
        // pot[14]: -0.5 au1 p_i.M.p_j over both orders -> -au1 M p_j
        const Dual e_pp = dexp(-(r2 * (iRpp * iRpp)));
        const Dual App1 = inv_r5 * (derf(rij * iRpp) - (2.0 * K_INV_SQRT_PI * iRpp) * rij * e_pp);
        const Dual Bpp1 = (4.0 * K_INV_SQRT_PI * iRpp3) * (inv_r * inv_r) * e_pp;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            gp[c] -= K_FQQ_AU1 * ((3.0 * rv[c] * pjr - r2 * pj[c]) * App1 - rv[c] * pjr * Bpp1);
    }
    gq = wave_sum_d(gq);
#pragma unroll
    for (int c = 0; c < 3; ++c) gp[c] = wave_sum_d(gp[c]);
    if (lane != 0) return;
    const double qi = x[3 * Nm + i];
    const double pi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    // self terms and the external field (pot[9])
    gq.v += K_FQQ * qi * (bc.inv_capac * K_NM2BOHR);
    gq.v -= bc.efield[0] * xi + bc.efield[1] * yi + bc.efield[2] * zi;
    gq.d -= bc.efield[0] * Gi[0] + bc.efield[1] * Gi[1] + bc.efield[2] * Gi[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gp[c].v += K_FQQ * pi[c] * (bc.inv_polar * K_NM2BOHR) - bc.efield[c];
        gE[3 * i + c] = gp[c].v;
        dgE[3 * i + c] = gp[c].d;
    }
    gE[3 * Nm + i] = gq.v;
    dgE[3 * Nm + i] = gq.d;
}

// Derivatives along one CPIM parameter (seed 0..3: alpha, c, R_q, R_p) at
// fixed positions and x: of the metal-metal force on metal i (mirroring the
// metal-metal terms of k_cpim_mm), of the CPIM energy (pair + self terms
// of pot[12..14]) and of the relay rows of i applied to x.  One wavefront per
// metal atom.
__global__ void __launch_bounds__(256)
k_theta_cpim(const double* __restrict__ r, BoxConst bc, const double* __restrict__ x, int seed,
             double alpha, double capac, double Rq, double Rp, double* __restrict__ dF,
             double* __restrict__ dE, double* __restrict__ dAx) {
    const int N = bc.N, Nm = bc.Nm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wave;
    if (i >= Nm) return;                               // no barrier in this kernel
    const Dual al = dl(alpha, seed == 0), cp = dl(capac, seed == 1);
    const Dual rq = dl(Rq, seed == 2), rp = dl(Rp, seed == 3);
    const Dual inv_polar = 1.0 / al, inv_capac = 1.0 / cp;
    // widths as in the reference: R_qq^2 = 2 R_q^2, R_pq^2 = R_p^2 + R_q^2, R_pp^2 = 2 R_p^2
    const Dual iRqq = 1.0 / dsqrt(rq * rq + rq * rq);
    const Dual iRpq = 1.0 / dsqrt(rp * rp + rq * rq);
    const Dual iRpp = 1.0 / dsqrt(rp * rp + rp * rp);
    const Dual iRpq3 = iRpq * iRpq * iRpq, iRpp3 = iRpp * iRpp * iRpp;
    const Dual iRpp5 = iRpp3 * iRpp * iRpp;
    const double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
    const double qi = x[3 * Nm + i];
    const double pi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    Dual f[3] = {dl(0.0), dl(0.0), dl(0.0)}, e = dl(0.0);
    Dual ax[4] = {dl(0.0), dl(0.0), dl(0.0), dl(0.0)};
    for (int j = lane; j < Nm; j += 64) {
        if (j == i) continue;
        double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
        pbc12(bc, dx, dy, dz);
        dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
        const double r2 = dist2_exact(dx, dy, dz);
        const double rij = sqrt(r2);
        const double inv_r = 1.0 / rij, inv_r2 = inv_r * inv_r;
        const double inv_r3 = inv_r2 * inv_r, inv_r5 = inv_r3 * inv_r2, inv_r7 = inv_r5 * inv_r2;
        const double qj = x[3 * Nm + j];
        const double pj[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
        const double rv[3] = {dx, dy, dz};
        Dual g[3] = {dl(0.0), dl(0.0), dl(0.0)};
        // T0
        const Dual e_qq = dexp(-r2 * (iRqq * iRqq));
        const Dual erf_qq = derf(rij * iRqq);
        const Dual A0 = inv_r3 * (erf_qq - 2.0 * K_INV_SQRT_PI * rij * (iRqq * e_qq));
        const Dual sqq = A0 * (K_FQQ_AU2 * qi * qj);
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] += sqq * rv[c];
        e += (0.5 * K_FQQ_AU1 * qi * qj * inv_r) * erf_qq;
        // T1
        const Dual e_pq = dexp(-r2 * (iRpq * iRpq));
        const Dual br_pq = derf(rij * iRpq) - 2.0 * K_INV_SQRT_PI * rij * (iRpq * e_pq);
        const Dual Apq0 = inv_r3 * br_pq;
        const Dual A1 = inv_r5 * br_pq;
        const Dual B1 = (4.0 * K_INV_SQRT_PI * inv_r2) * (iRpq3 * e_pq);
        const double w[3] = {qj * pi[0] - qi * pj[0], qj * pi[1] - qi * pj[1], qj * pi[2] - qi * pj[2]};
        const double wr = w[0] * dx + w[1] * dy + w[2] * dz;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            g[b] += (A1 * (3.0 * wr * rv[b] - w[b] * r2) - B1 * (wr * rv[b])) * K_FQQ_AU2;
        const double pjr = pj[0] * dx + pj[1] * dy + pj[2] * dz;
        const double pir = pi[0] * dx + pi[1] * dy + pi[2] * dz;
        e += (0.5 * K_FQQ_AU1 * (qj * pir - qi * pjr)) * Apq0;
        // T2
        const Dual e_pp = dexp(-r2 * (iRpp * iRpp));
        const Dual erf_pp = derf(rij * iRpp);
        const Dual App1 = inv_r5 * (erf_pp - 2.0 * K_INV_SQRT_PI * rij * (iRpp * e_pp));
        const Dual Bpp1 = (4.0 * K_INV_SQRT_PI * inv_r2) * (iRpp3 * e_pp);
        const Dual A2 = inv_r7 * (3.0 * erf_pp - 6.0 * K_INV_SQRT_PI * rij * (iRpp * e_pp) -
                                  4.0 * r2 * rij * K_INV_SQRT_PI * (iRpp3 * e_pp));
        const Dual B2 = (8.0 * K_INV_SQRT_PI * inv_r2) * (iRpp5 * e_pp);
        const double pipj = pi[0] * pj[0] + pi[1] * pj[1] + pi[2] * pj[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double t5 = 5.0 * pir * pjr * rv[c] - r2 * (pir * pj[c] + pjr * pi[c] + pipj * rv[c]);
            g[c] -= (A2 * t5 - B2 * (pir * pjr * rv[c])) * K_FQQ_AU2;
        }
        e -= 0.5 * K_FQQ_AU1 * ((3.0 * pir * pjr - pipj * r2) * App1 - (pir * pjr) * Bpp1);
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] -= g[c];
        // relay block (i, j) applied to x_j: c1 = App1, c2 = Bpp1, c3 = Apq0,
        // cq = -erf(r/R_qq)/r (the entries of cpff.c:326-428)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            ax[a] += Apq0 * (rv[a] * qj) - (App1 * (3.0 * rv[a] * pjr - r2 * pj[a]) - Bpp1 * (rv[a] * pjr));
        ax[3] += -(Apq0 * pjr) - erf_qq * (inv_r * qj);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = wave_sum_d(f[c]);
    e = wave_sum_d(e);
#pragma unroll
    for (int c = 0; c < 4; ++c) ax[c] = wave_sum_d(ax[c]);
    if (lane != 0) return;
    // self terms and the diagonal blocks (1/alpha, -1/c)
    e += (0.5 * K_FQQ * qi * qi * K_NM2BOHR) * inv_capac;
    e += (0.5 * K_FQQ * (pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2]) * K_NM2BOHR) * inv_polar;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ax[a] += inv_polar * pi[a];
        dAx[3 * i + a] = ax[a].d;
        dF[a * Nm + i] = f[a].d;
    }
    ax[3] -= inv_capac * qi;
    dAx[3 * Nm + i] = ax[3].d;
    dE[i] = e.d;
}

__device__ __forceinline__ double ipow_d(double x, int n) {
    double r = 1.0;
    for (int k = 0; k < n; ++k) r *= x;
    return r;
}

// Derivatives along one Sutton-Chen parameter (seed 4: eps, 5: c) of the
// Sutton-Chen force on metal i and of its energy terms (force.c:468-542);
// inv_sqrt_dens from the density pass of the same force call.
__global__ void __launch_bounds__(256)
k_theta_sc(const double* __restrict__ r, BoxConst bc, const int* __restrict__ n_back,
           const int* __restrict__ nbr, int stride, const double* __restrict__ isd, int seed,
           double* __restrict__ dF, double* __restrict__ dE) {
    const int N = bc.N, Nm = bc.Nm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wave;
    if (i >= Nm) return;                               // no barrier in this kernel
    const Dual eps = dl(bc.qsc_eps, seed == 4), csc = dl(bc.qsc_c, seed == 5);
    const double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
    const double isd_i = isd[i];
    const Dual cm_2 = csc * (bc.qsc_m / 2.0);
    const int nb = n_back[i];
    const int* tail = nbr + (size_t)i * stride + stride - nb;
    Dual f[3] = {dl(0.0), dl(0.0), dl(0.0)}, e = dl(0.0);
    for (int k = lane; k < nb; k += 64) {
        const int j = tail[k];
        double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
        pbc12(bc, dx, dy, dz);
        const double r2 = dist2_exact(dx, dy, dz);
        if (!(r2 < bc.rc2)) continue;
        const double rij = sqrt(r2);
        const double a_r = bc.qsc_a / rij;
        const double a_n = ipow_d(a_r, bc.qsc_n), a_m = ipow_d(a_r, bc.qsc_m);
        const Dual fij = eps * (bc.qsc_n * a_n - cm_2 * ((isd_i + isd[j]) * a_m)) / r2;
        f[0] -= fij * dx; f[1] -= fij * dy; f[2] -= fij * dz;
        e += eps * (0.5 * a_n);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = wave_sum_d(f[c]);
    e = wave_sum_d(e);
    if (lane != 0) return;
    e -= (eps * csc) / isd_i;                          // embedding term of atom i
#pragma unroll
    for (int c = 0; c < 3; ++c) dF[c * Nm + i] = f[c].d;
    dE[i] = e.d;
}

}  // namespace

// ------------------------------------------------------------------ host

struct FitBuffers {
    double *G, *gE, *dgE, *rhs, *lam, *dF, *dE, *dAx;
};

static int fit_buffers(capac_ctx* c) {
    if (c->fit) return 0;
    FitBuffers* fb = (FitBuffers*)calloc(1, sizeof(FitBuffers));
    if (!fb) { capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
    c->fit = fb;
    const size_t nm = (size_t)c->bc.Nm, n = (size_t)c->n_mat;
    double** p[8] = {&fb->G, &fb->gE, &fb->dgE, &fb->rhs, &fb->lam, &fb->dF, &fb->dE, &fb->dAx};
    const size_t sz[8] = {3 * nm, n, n, n, n, 3 * nm, nm, n};
    for (int k = 0; k < 8; ++k) {
        CHECK_HIP(hipMalloc(p[k], sz[k] * sizeof(double)));
        CHECK_HIP(hipMemset(*p[k], 0, sz[k] * sizeof(double)));
    }
    return 0;
}

void fit_destroy(capac_ctx* c) {
    FitBuffers* fb = (FitBuffers*)c->fit;
    if (!fb) return;
    double* p[8] = {fb->G, fb->gE, fb->dgE, fb->rhs, fb->lam, fb->dF, fb->dE, fb->dAx};
    for (int k = 0; k < 8; ++k)
        if (p[k]) (void)hipFree(p[k]);
    free(fb);
    c->fit = NULL;
}

// apply theta to the device constants (widths as free parameters)
void fit_apply_theta(capac_ctx* c) {
    BoxConst& b = c->bc;
    const double* t = c->theta;
    b.inv_polar = 1.0 / t[0];
    b.inv_capac = 1.0 / t[1];
    b.inv_R_qq = 1.0 / sqrt(t[2] * t[2] + t[2] * t[2]);
    b.inv_R_pq = 1.0 / sqrt(t[3] * t[3] + t[2] * t[2]);
    b.inv_R_pp = 1.0 / sqrt(t[3] * t[3] + t[3] * t[3]);
    b.qsc_eps = t[4];
    b.qsc_c = t[5];
}

static int d2h(capac_ctx* c, double* dst, const double* src, size_t n) {
    CHECK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int fit_frame_loss_grad(capac_ctx* c, const double* f_ref, double e_ref, double w_e, double* loss,
                        double grad[CAPAC_NTHETA]) {
    const int Nm = c->bc.Nm, N = c->bc.N, n = c->n_mat;
    if (Nm == 0 || n == 0) { capac_set_error("the fitting loss needs metal atoms with CPIM"); return CAPAC_ERR_ARG; }
    int rc = fit_buffers(c);
    if (rc) return rc;
    FitBuffers* fb = (FitBuffers*)c->fit;
    double pot[CAPAC_NPOT], vir[9];
    rc = capac_force(c, pot, vir, NULL);
    if (rc) return rc;
    double* f = (double*)malloc(sizeof(double) * 3 * (size_t)N);
    double* h = (double*)malloc(sizeof(double) * (3 * (size_t)Nm + 4 * (size_t)n + Nm));
    if (!f || !h) { free(f); free(h); capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
    double* G = h;                      // 3 Nm
    double* v1 = G + 3 * Nm;            // n
    double* v2 = v1 + n;                // n
    double* v3 = v2 + n;                // n
    double* v4 = v3 + n;                // n
    double* ve = v4 + n;                // Nm
#define FIT_TRY(x) do { rc = (x); if (rc) { free(f); free(h); return rc; } } while (0)
    FIT_TRY(d2h(c, f, c->f, 3 * (size_t)N));
    double E = 0.0;
    for (int k = 1; k < CAPAC_NPOT; ++k) E += pot[k];
    double L = 0.0;
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < Nm; ++i) {
            double df = f[(size_t)d * N + i] - f_ref[(size_t)d * Nm + i];
            L += df * df;
            G[(size_t)d * Nm + i] = 2.0 * df;
        }
    const double dE = E - e_ref;
    L += w_e * dE * dE;
    if (loss) *loss = L;
    if (!grad) { free(f); free(h); return 0; }
    // adjoint right-hand side dL/dx = -D_G (dE/dx) + 2 w_E (E - E_ref) dE/dx
    if (hipMemcpyAsync(fb->G, G, sizeof(double) * 3 * Nm, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        free(f); free(h); capac_set_error("copy failed"); return CAPAC_ERR_HIP;
    }
    const int nblk = (Nm + 3) / 4;
    k_gradx<<<nblk, 256, 0, c->stream>>>(c->r, c->charge, c->bc, c->n_front, c->nbr, c->nb_stride, c->x,
                                         fb->G, fb->gE, fb->dgE);
    FIT_TRY(hipGetLastError() == hipSuccess ? 0 : CAPAC_ERR_HIP);
    FIT_TRY(d2h(c, v1, fb->gE, 4 * (size_t)Nm));
    FIT_TRY(d2h(c, v2, fb->dgE, 4 * (size_t)Nm));
    for (int k = 0; k < n; ++k) v3[k] = k < 4 * Nm ? -v2[k] + 2.0 * w_e * dE * v1[k] : 0.0;
    for (int k = 0; k < n; ++k) v4[k] = 0.0;
    if (hipMemcpyAsync(fb->rhs, v3, sizeof(double) * n, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(fb->lam, v4, sizeof(double) * n, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        free(f); free(h); capac_set_error("copy failed"); return CAPAC_ERR_HIP;
    }
    // the relay is symmetric: A lambda = dL/dx with the same solver
    const int iters = c->solve_iters, kind = c->solve_kind;
    const double rtr = c->solve_rTr, btb = c->solve_bTb;
    rc = solver_run(c, fb->rhs, fb->lam);
    c->fit_adjoint_iters = c->solve_iters;
    c->solve_iters = iters; c->solve_kind = kind; c->solve_rTr = rtr; c->solve_bTb = btb;
    FIT_TRY(rc);
    FIT_TRY(d2h(c, v4, fb->lam, (size_t)n));          // lambda
    for (int s = 0; s < CAPAC_NTHETA; ++s) {
        if (s < 4) {
            k_theta_cpim<<<nblk, 256, 0, c->stream>>>(c->r, c->bc, c->x, s, c->theta[0], c->theta[1],
                                                      c->theta[2], c->theta[3], fb->dF, fb->dE, fb->dAx);
        } else {
            k_theta_sc<<<nblk, 256, 0, c->stream>>>(c->r, c->bc, c->n_back, c->nbr, c->nb_stride, c->isd, s,
                                                    fb->dF, fb->dE);
        }
        FIT_TRY(hipGetLastError() == hipSuccess ? 0 : CAPAC_ERR_HIP);
        FIT_TRY(d2h(c, v1, fb->dF, 3 * (size_t)Nm));
        FIT_TRY(d2h(c, ve, fb->dE, (size_t)Nm));
        double g = 0.0, de = 0.0;
        for (int k = 0; k < 3 * Nm; ++k) g += G[k] * v1[k];
        for (int k = 0; k < Nm; ++k) de += ve[k];
        g += 2.0 * w_e * dE * de;
        if (s < 4) {
            FIT_TRY(d2h(c, v2, fb->dAx, 4 * (size_t)Nm));
            double la = 0.0;
            for (int k = 0; k < 4 * Nm; ++k) la += v4[k] * v2[k];
            g -= la;
        }
        grad[s] = g;
    }
#undef FIT_TRY
    free(f);
    free(h);
    return 0;
}
