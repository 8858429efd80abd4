// Energy minimisation by steepest descent or conjugate gradients, the
// reference's run_type em / cg loop body (mpi-main.c:706-865) on the device:
//   move   r_old = r; r += s / F_max * length (fixed metal stays);
//          then (in the caller) RATTLE_1 on the positions and v = 0
//   force  one force call; F_max and F_rms
//   dir    s = F + gamma s, F_old = F, gamma = 0 (em) or
//          (F.F - F_old.F) / (F_old.F_old) (cg, mpi-main.c:826-843)
// The host keeps the step-length rule and the convergence test.
#include "capac_internal.h"

// Minimiser arithmetic is evaluated as written (no contraction), as in the
// rest of the integrator.
#pragma clang fp contract(off)

namespace {

__global__ void __launch_bounds__(256)
k_em_move(int N, int Nm, int fix_metal, double length, const double* __restrict__ md,
          const double* __restrict__ s, double* __restrict__ r, double* __restrict__ old_r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double fmax = md[23];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t k = (size_t)c * N + i;
        old_r[k] = r[k];
        if (fix_metal == 1 && i < Nm) continue;
        r[k] += s[k] / fmax * length;
    }
}

// F.F, F_old.F, F_old.F_old per block (fixed order within the block)
__global__ void __launch_bounds__(256)
k_em_dots(int n3, const double* __restrict__ f, const double* __restrict__ f_old, double* __restrict__ part) {
    __shared__ double s_w[4][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    if (k < n3) {
        const double fk = f[k], gk = f_old[k];
        a = fk * fk;
        b = gk * fk;
        c = gk * gk;
    }
    a = wave_sum(a); b = wave_sum(b); c = wave_sum(c);
    if (lane == 0) { s_w[wave][0] = a; s_w[wave][1] = b; s_w[wave][2] = c; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += s_w[w][threadIdx.x];
        part[3 * blockIdx.x + threadIdx.x] = t;
    }
}

// gamma from the block partials (one workgroup, fixed order), then the new
// direction and F_old
__global__ void __launch_bounds__(1024)
k_em_gamma(int nblk, const double* __restrict__ part, int cg, double* __restrict__ gamma_out) {
    __shared__ double s_w[16][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
        a += part[3 * k];
        b += part[3 * k + 1];
        c += part[3 * k + 2];
    }
    a = wave_sum(a); b = wave_sum(b); c = wave_sum(c);
    if (lane == 0) { s_w[wave][0] = a; s_w[wave][1] = b; s_w[wave][2] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double g22 = 0.0, g12 = 0.0, g11 = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { g22 += s_w[w][0]; g12 += s_w[w][1]; g11 += s_w[w][2]; }
        gamma_out[0] = cg ? (g22 - g12) / g11 : 0.0;
    }
}

__global__ void __launch_bounds__(256)
k_em_dir(int n3, const double* __restrict__ gamma, const double* __restrict__ f, double* __restrict__ s,
         double* __restrict__ f_old) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    const double g = gamma[0];
    const double fk = f[k];
    s[k] = fk + g * s[k];
    f_old[k] = fk;
}

}  // namespace

struct EmState {
    double *s, *f_old, *part, *gamma;
    int nblk;
};

void em_destroy(capac_ctx* c) {
    EmState* e = (EmState*)c->em;
    if (!e) return;
    double* p[4] = {e->s, e->f_old, e->part, e->gamma};
    for (int k = 0; k < 4; ++k)
        if (p[k]) (void)hipFree(p[k]);
    free(e);
    c->em = NULL;
}

// s = F, F_old = F, v = 0 after the first force call (mpi-main.c:716-724)
int em_begin(capac_ctx* c) {
    const size_t n3 = 3 * (size_t)c->bc.N;
    if (!c->em) {
        EmState* e = (EmState*)calloc(1, sizeof(EmState));
        if (!e) { capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
        c->em = e;
        e->nblk = (int)((n3 + 255) / 256);
        CHECK_HIP(hipMalloc(&e->s, n3 * sizeof(double)));
        CHECK_HIP(hipMalloc(&e->f_old, n3 * sizeof(double)));
        CHECK_HIP(hipMalloc(&e->part, 3 * (size_t)e->nblk * sizeof(double)));
        CHECK_HIP(hipMalloc(&e->gamma, sizeof(double)));
    }
    EmState* e = (EmState*)c->em;
    CHECK_HIP(hipMemcpyAsync(e->s, c->f, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    CHECK_HIP(hipMemcpyAsync(e->f_old, c->f, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return md_fmax(c);
}

int em_move(capac_ctx* c, double length) {
    EmState* e = (EmState*)c->em;
    if (!e) { capac_set_error("capac_em_begin has not been called"); return CAPAC_ERR_ARG; }
    const int N = c->bc.N;
    k_em_move<<<(N + 255) / 256, 256, 0, c->stream>>>(N, c->bc.Nm, c->fix_metal, length, c->md, e->s, c->r,
                                                      c->old_r);
    CHECK_HIP(hipGetLastError());
    return 0;
}

int em_direction(capac_ctx* c, int cg) {
    EmState* e = (EmState*)c->em;
    if (!e) { capac_set_error("capac_em_begin has not been called"); return CAPAC_ERR_ARG; }
    const int n3 = 3 * c->bc.N;
    k_em_dots<<<e->nblk, 256, 0, c->stream>>>(n3, c->f, e->f_old, e->part);
    k_em_gamma<<<1, 1024, 0, c->stream>>>(e->nblk, e->part, cg, e->gamma);
    k_em_dir<<<e->nblk, 256, 0, c->stream>>>(n3, e->gamma, c->f, e->s, e->f_old);
    CHECK_HIP(hipGetLastError());
    return 0;
}
