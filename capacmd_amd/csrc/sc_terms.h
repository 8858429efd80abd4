// Sutton-Chen passes as workgroup bodies (sc.hip's launches and the merged
// launches of cpim.hip); see sc.hip for the terms.
#pragma once
#include "capac_internal.h"


__device__ __forceinline__ double ipow(double x, int n) {
    double r = 1.0;
    while (n > 0) {
        if (n & 1) r *= x;
        x *= x;
        n >>= 1;
    }
    return r;
}

// pass 1 for the four metal atoms of workgroup b (one per wavefront)
__device__ __forceinline__ void sc_dens_block(int b, const double* __restrict__ r, const BoxConst& bc,
                                              const int* __restrict__ n_back, const int* __restrict__ nbr, int stride,
                                              double* __restrict__ isd) {
    const int N = bc.N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int i = b * 4 + wave;
    if (i >= bc.Nm) return;
    double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
    int nb = n_back[i];
    const int* tail = nbr + (size_t)i * stride + stride - nb;
    double dens = 0.0;
    for (int k = lane; k < nb; k += 64) {
        int j = tail[k];
        double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
        pbc12(bc, dx, dy, dz);
        double r2 = dist2_exact(dx, dy, dz);
        if (!(r2 < bc.rc2)) continue;         // rows reach rc + skin
        dens += ipow(bc.qsc_a / sqrt(r2), bc.qsc_m);
    }
    dens = wave_sum(dens);
    if (lane == 0) isd[i] = 1.0 / sqrt(dens);
}

// pass 2 for the four metal atoms of workgroup b; its energy and virial
// partials to partial[b]; the force written to (or, add, added to) f with
// component stride f_stride
__device__ __forceinline__ void sc_force_block(int b, const double* __restrict__ r, const BoxConst& bc,
                                               const int* __restrict__ n_back, const int* __restrict__ nbr,
                                               int stride, const double* __restrict__ isd, double* __restrict__ f,
                                               int f_stride, int add, double* __restrict__ partial,
                                               double (*s_red)[10]) {
    const int N = bc.N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int i = b * 4 + wave;
    double e1 = 0.0, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < bc.Nm) {
        double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
        double isd_i = isd[i];
        double cm_2 = bc.qsc_c * bc.qsc_m / 2.0;
        int nb = n_back[i];
        const int* tail = nbr + (size_t)i * stride + stride - nb;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int k = lane; k < nb; k += 64) {
            int j = tail[k];
            double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
            pbc12(bc, dx, dy, dz);
            double r2 = dist2_exact(dx, dy, dz);
            if (!(r2 < bc.rc2)) continue;     // rows reach rc + skin
            double rij = sqrt(r2);
            double a_r = bc.qsc_a / rij;
            double a_n = ipow(a_r, bc.qsc_n);
            double a_m = ipow(a_r, bc.qsc_m);
            double fij = bc.qsc_eps * (bc.qsc_n * a_n - cm_2 * (isd_i + isd[j]) * a_m) / r2;
            double gx = fij * dx, gy = fij * dy, gz = fij * dz;
            fx -= gx; fy -= gy; fz -= gz;
            v[0] -= 0.5 * dx * gx; v[1] -= 0.5 * dx * gy; v[2] -= 0.5 * dx * gz;
            v[3] -= 0.5 * dy * gx; v[4] -= 0.5 * dy * gy; v[5] -= 0.5 * dy * gz;
            v[6] -= 0.5 * dz * gx; v[7] -= 0.5 * dz * gy; v[8] -= 0.5 * dz * gz;
            e1 += 0.5 * bc.qsc_eps * a_n;
        }
        fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
        if (lane == 0) {
            if (add) { f[i] += fx; f[f_stride + i] += fy; f[2 * f_stride + i] += fz; }
            else { f[i] = fx; f[f_stride + i] = fy; f[2 * f_stride + i] = fz; }
            e1 -= bc.qsc_eps * bc.qsc_c / isd_i;      // embedding term of atom i
        }
    }
    double vals[10] = {e1, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
#pragma unroll
    for (int c = 0; c < 10; ++c) {
        double s = wave_sum(vals[c]);
        if (lane == 0) s_red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double s = 0.0;
        if (threadIdx.x < 10)
            for (int w = 0; w < 4; ++w) s += s_red[w][threadIdx.x];
        partial[b * 16 + threadIdx.x] = s;
    }
}

