// Block-Jacobi preconditioner of the FP32-refined path's GMRES.
//
// The metal atoms are cut into the relay's blocks of 64 (relay_tile.h); the
// 256 unknowns of block b (dipoles of its 64 atoms, then their charges) and
// the relay rows and columns between them form the diagonal block A_bb:
//     rho < 192: dipole component rho % 3 of atom 64 b + rho / 3
//     rho >= 192: charge of atom 64 b + rho - 192
// (unknowns past Nm stand for an identity row).  The preconditioner applies
// A_bb^-1 to the block's unknowns and 1 to the Lagrange rows.  The signed
// Jacobi K of the reference (its diagonal) is the same thing with blocks of
// one unknown; with blocks of 64 atoms the Au2057 relay needs 16 GMRES
// products to the reference's stop test against 25 with K (dense oracle
// relay, quadratic warm start), and an inverse formed 20 MD steps earlier
// serves as well as a fresh one, so the blocks are inverted every few
// assemblies only.
//
// A_bb is symmetric quasi-definite: its dipole block is positive definite
// (1/alpha on the diagonal) and its charge block negative definite (-1/c),
// so Gauss-Jordan elimination needs no pivoting (every leading principal
// submatrix is nonsingular).  One 1024-thread workgroup per block keeps the
// 256 x 256 FP32 matrix in registers and inverts it in place (gj256).
//
// Coarse space.  Block Jacobi leaves the solve's slowest modes: charge
// moving between blocks.  In MD they are also what a solve at the
// reference's tolerance leaves behind, and the warm-start extrapolation
// carries that error into the next step's residual, so with block Jacobi
// alone the Au2057 steps still take ~25 products.  An additive coarse
// correction over four charge vectors per block -- 1 and the atoms' x, y, z
// in the block, orthonormalised -- takes them out: M^-1 = B^-1 + Z Ac^-1 Z^T,
// Ac = Z^T A Z (4 Nb unknowns, negative definite like the charge block).
// On recorded Au2057 MD frames the solve chain (each step from the
// extrapolation of the chain's own solutions) falls from 26 products per
// step (Jacobi) and 21 (block Jacobi) to 14 (dense FP64 relay).
//
// Stored inverse (FP32): element (rho, sigma) of block b at
//     inv[b * 65536 + (sigma / 4) * 1024 + rho * 4 + sigma % 4],
// so the 64 lanes of an application, one per output row, read 64 consecutive
// float4 (four columns each) per load.
#include <stdlib.h>

#include "capac_internal.h"
#include "precond.h"
#include "relay.h"
#include "relay_tile.h"

#ifndef CAPAC_BJ_COARSE
#define CAPAC_BJ_COARSE 1
#endif

namespace {

constexpr int kBjThreads = 1024;

// the four FP32 coefficients of ordered metal pair (i, j), i != j
__device__ __forceinline__ float4 pair_coef(const float* __restrict__ relay, int tiles, int Nb, int ld, int i,
                                            int j) {
    if (tiles) return reinterpret_cast<const float4*>(relay)[tile_slot(i, j, Nb)];
    const float* p = relay + (size_t)i * kRelay32Planes * ld + j;
    return make_float4(p[0], p[ld], p[2 * (size_t)ld], p[3 * (size_t)ld]);
}

// In-place Gauss-Jordan inverse of the 256 x 256 FP32 matrix that a
// 1024-thread workgroup holds in registers, thread (ty, tx) = (t / 32, t % 32)
// holding rows ty + 32 a and columns tx + 32 c of A[a][c] (a, c = 0..7);
// no pivoting (the callers' matrices are quasi-definite).  Pivot
// k = 32 cb + m sits in register column cb of the threads with tx = m, a
// compile-time index when the loop over cb is unrolled.  Step k: the threads
// holding column k and row k post them to LDS, column k becomes e_k, and
// every element takes a_ij -= u_i v_j with u = column k (u_k = a_kk - 1) and
// v = row k / a_kk (v_k = 1 / a_kk): the Gauss-Jordan update in one
// rank-one form.  The LDS rows are double-buffered: one barrier per step.
// Only pivots k < nk are taken: rows and columns from nk on must be those of
// the identity (a padding), which the remaining steps would leave as they are.
__device__ __forceinline__ void gj256(float (&A)[8][8], float (*s_col)[kBjUnk], float (*s_row)[kBjUnk], int tx,
                                      int ty, int nk) {
#pragma unroll
    for (int cb = 0; cb < 8; ++cb) {
        for (int m = 0; m < 32; ++m) {
            const int k = 32 * cb + m, buf = k & 1;
            if (k >= nk) break;
            if (tx == m) {
#pragma unroll
                for (int a = 0; a < 8; ++a) s_col[buf][ty + 32 * a] = A[a][cb];
            }
            if (ty == m) {
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) s_row[buf][tx + 32 * cc] = A[cb][cc];
            }
            if (tx == m) {
#pragma unroll
                for (int a = 0; a < 8; ++a) A[a][cb] = (a == cb && ty == m) ? 1.0f : 0.0f;
            }
            __syncthreads();
            const float p = s_row[buf][k];
            const float ip = 1.0f / p;
            float u[8], v[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) u[a] = (a == cb && ty == m) ? p - 1.0f : s_col[buf][ty + 32 * a];
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) v[cc] = (cc == cb && tx == m) ? ip : s_row[buf][tx + 32 * cc] * ip;
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) A[a][cc] = fmaf(-u[a], v[cc], A[a][cc]);
        }
    }
}

__global__ void __launch_bounds__(kBjThreads)
k_bj_build(const float* __restrict__ relay, int tiles, int Nb, int ld, const double* __restrict__ r,
           const double* __restrict__ mu, BoxConst bc, float* __restrict__ inv, double* __restrict__ r_at) {
    main_stream_priority();
    __shared__ float s_col[2][kBjUnk], s_row[2][kBjUnk];
    const int b = blockIdx.x, Nm = bc.Nm;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (threadIdx.x < kTileEdge && b * kTileEdge + (int)threadIdx.x < Nm) {
        // the block's metal positions, kept with its inverse
        const int gi = b * kTileEdge + threadIdx.x;
        r_at[gi] = r[gi];
        r_at[Nm + gi] = r[bc.N + gi];
        r_at[2 * Nm + gi] = r[2 * bc.N + gi];
    }
    const bool compact = mu[4 * Nm] != 0.0;
    float A[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int rho = ty + 32 * a;
        const int ai = rho < 3 * kTileEdge ? rho / 3 : rho - 3 * kTileEdge;
        const int ci = rho < 3 * kTileEdge ? rho % 3 : 3;
        const int gi = b * kTileEdge + ai;
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int sig = tx + 32 * cc;
            const int aj = sig < 3 * kTileEdge ? sig / 3 : sig - 3 * kTileEdge;
            const int cj = sig < 3 * kTileEdge ? sig % 3 : 3;
            const int gj = b * kTileEdge + aj;
            float v;
            if (gi >= Nm || gj >= Nm) {
                v = rho == sig ? 1.0f : 0.0f;
            } else if (ai == aj) {
                v = rho != sig ? 0.0f : (float)(ci < 3 ? bc.inv_polar : -bc.inv_capac);
            } else {
                double d[3];
                if (compact) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) d[k] = mu[4 * gj + k] - mu[4 * gi + k];
                } else {
                    metal_sep(r, bc, gi, gj, d[0], d[1], d[2]);
                }
                const float4 cf = pair_coef(relay, tiles, Nb, ld, gi, gj);
                // block (i, j) of relay.h: [g_a r r^T + g_b 1, c3 r; -c3 r^T, cq]
                double e;
                if (ci < 3 && cj < 3) e = cf.x * d[ci] * d[cj] + (ci == cj ? cf.y : 0.0f);
                else if (ci < 3) e = cf.z * d[ci];
                else if (cj < 3) e = -cf.z * d[cj];
                else e = cf.w;
                v = (float)e;
            }
            A[a][cc] = v;
        }
    }
    gj256(A, s_col, s_row, tx, ty, kBjUnk);
    float* dst = inv + (size_t)b * kBjUnk * kBjUnk;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int rho = ty + 32 * a, sig = tx + 32 * cc;
            dst[(sig >> 2) * (4 * kBjUnk) + rho * 4 + (sig & 3)] = A[a][cc];
        }
}

// out = M q (Add = false) or out += M q (Add = true) for the preconditioner
// M: workgroup (b, s), s = 0..3, forms the 64 unknowns 64 s .. 64 s + 63 of
// block b; its 256 threads take four column quarters of each row, added in
// order.  A last workgroup copies (or adds) the Lagrange rows.  Guard 1: no-op
// once the solve is done; guard 2: no-op until it is.
template <bool Add>
__global__ void __launch_bounds__(256)
k_bj_apply(int Nm, int Nb, int n_mat, const float* __restrict__ inv, const double* __restrict__ q,
           double* __restrict__ out, const float4* __restrict__ cw, double* __restrict__ cs,
           const SolveScalars* __restrict__ sc, int guard) {
    main_stream_priority();
    __shared__ double s_q[kBjUnk];
    __shared__ double s_red[4][kTileEdge];
    if (guard == 1 && sc->done) return;
    if (guard == 2 && !sc->done) return;
    const int b = blockIdx.x >> 2, s = blockIdx.x & 3, t = threadIdx.x;
    if (b >= Nb) {
        for (int k = 4 * Nm + t; k < n_mat; k += blockDim.x) out[k] = Add ? out[k] + q[k] : q[k];
        return;
    }
    const int ro = s * kTileEdge + (t & 63), ch = t >> 6;
    // this thread's 16 float4 of the matrix, issued before the vector's staging
    const float4* mp = reinterpret_cast<const float4*>(inv + (size_t)b * kBjUnk * kBjUnk) + ro;
    float4 mv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) mv[i] = mp[(size_t)(16 * ch + i) * kBjUnk];
    {
        const int row = bj_row(b, t, Nm);
        s_q[t] = row >= 0 ? q[row] : 0.0;
    }
    __syncthreads();
    if (cs && s == 3 && t >= 3 * kTileEdge) {
        // the block's coarse components Z_b^T q (the charges are the last 64 unknowns)
        const float4 z = cw[b * kTileEdge + (t & 63)];
        const double qv = s_q[t];
        const double c0 = wave_sum(z.x * qv), c1 = wave_sum(z.y * qv), c2 = wave_sum(z.z * qv),
                     c3 = wave_sum(z.w * qv);
        if ((t & 63) == 0) {
            cs[4 * b] = c0; cs[4 * b + 1] = c1; cs[4 * b + 2] = c2; cs[4 * b + 3] = c3;
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int c0 = 4 * (16 * ch + i);
        acc += (double)mv[i].x * s_q[c0] + (double)mv[i].y * s_q[c0 + 1] + (double)mv[i].z * s_q[c0 + 2] +
               (double)mv[i].w * s_q[c0 + 3];
    }
    s_red[ch][t & 63] = acc;
    __syncthreads();
    if (t >= kTileEdge) return;
    const int row = bj_row(b, ro, Nm);
    if (row < 0) return;
    const double v = ((s_red[0][t] + s_red[1][t]) + s_red[2][t]) + s_red[3][t];
    out[row] = Add ? out[row] + v : v;
}

// ---- coarse space ----

// per block (one wavefront): the coarse basis over its charges, the vectors
// 1, x, y, z (separations from the block's first atom) orthonormalised by
// modified Gram-Schmidt; one that depends on the earlier ones (what is left
// of it under 1e-6 of its norm: a block of one atom, a planar block) is
// dropped, its entries zero and act[4 b + k] = 0.  w[i] holds atom i's four
// entries.
__global__ void __launch_bounds__(64)
k_cg_basis(const double* __restrict__ r, BoxConst bc, float4* __restrict__ w, int* __restrict__ act) {
    const int b = blockIdx.x, l = threadIdx.x, i = b * kTileEdge + l;
    const bool valid = i < bc.Nm;
    double z[4] = {valid ? 1.0 : 0.0, 0.0, 0.0, 0.0};
    if (valid) metal_sep(r, bc, b * kTileEdge, i, z[1], z[2], z[3]);
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double x = z[k];
        const double n0 = sqrt(wave_sum(x * x));
#pragma unroll
        for (int m = 0; m < k; ++m) x -= wave_sum(v[m] * x) * v[m];
        const double n1 = sqrt(wave_sum(x * x));
        const bool on = n0 > 0.0 && n1 > 1.0e-6 * n0;
        v[k] = on ? x / n1 : 0.0;
        if (l == 0) act[4 * b + k] = on;
    }
    w[i] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}

// Ac = Z^T A Z from the relay's charge-charge coefficients cq: one
// workgroup per block pair (I <= J, numbered as the tiles), the 4 x 4 block
// sum_{i in I, j in J} w(i) cq(i, j) w(j)^T over the pair slots, and on a
// diagonal pair the diagonal -1/c; written to (I, J) and (J, I).  A dropped
// basis vector keeps a 1 on the diagonal (its row and column are zero).
__global__ void __launch_bounds__(256)
k_cg_coarse(const float* __restrict__ relay, int tiles, int Nb, int ld, const float4* __restrict__ w,
            const int* __restrict__ act, BoxConst bc, double* __restrict__ A, int nc) {
    __shared__ double s_w[4][16];
    const int Nm = bc.Nm, t = threadIdx.x;
    int I, J;
    tile_coords(blockIdx.x, Nb, I, J);
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    for (int e = t; e < kTileSlots; e += blockDim.x) {
        const int k = e / kTileEdge, l = e % kTileEdge;
        const int i = I * kTileEdge + l, j = J * kTileEdge + ((l + k) & (kTileEdge - 1));
        if (i == j || i >= Nm || j >= Nm) continue;
        const double cq = pair_coef(relay, tiles, Nb, ld, i, j).w;
        const float4 a = w[i], c = w[j];
        const double wa[4] = {a.x * cq, a.y * cq, a.z * cq, a.w * cq}, wc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[4 * p + q] += wa[p] * wc[q];
    }
    if (I == J && t < kTileEdge && I * kTileEdge + t < Nm) {
        const float4 a = w[I * kTileEdge + t];
        const double wa[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[4 * p + q] -= bc.inv_capac * wa[p] * wa[q];
    }
    const int wave = t >> 6, lane = t & 63;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const double v = wave_sum(acc[e]);
        if (lane == 0) s_w[wave][e] = v;
    }
    __syncthreads();
    if (t >= 16) return;
    const int p = t >> 2, q = t & 3;
    double v = ((s_w[0][t] + s_w[1][t]) + s_w[2][t]) + s_w[3][t];
    if (I == J && p == q && !act[4 * I + p]) v = 1.0;
    A[(size_t)(4 * I + p) * nc + 4 * J + q] = v;
    if (I != J) A[(size_t)(4 * J + q) * nc + 4 * I + p] = v;
}

// Ac^-1 in place (negative definite: no pivoting), one 1024-thread
// workgroup: Ac in FP32 registers, padded to 256 with identity rows, gj256,
// written back in FP64 (a preconditioner needs no more than FP32)
__global__ void __launch_bounds__(kBjThreads)
k_cg_invert(double* __restrict__ A, int nc) {
    main_stream_priority();
    __shared__ float s_col[2][kBjUnk], s_row[2][kBjUnk];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    float M[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int i = ty + 32 * a, j = tx + 32 * cc;
            M[a][cc] = i < nc && j < nc ? (float)A[(size_t)i * nc + j] : (i == j ? 1.0f : 0.0f);
        }
    gj256(M, s_col, s_row, tx, ty, nc);
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int i = ty + 32 * a, j = tx + 32 * cc;
            if (i < nc && j < nc) A[(size_t)i * nc + j] = M[a][cc];
        }
}

// out (charges) += Z Ac^-1 s: workgroup b (one wavefront) forms its block's
// four coarse values c = rows 4b..4b+3 of Ac^-1 times s, then adds
// w(i) . c to the charge of each of its atoms.  Guards as k_bj_apply.
__global__ void __launch_bounds__(64)
k_cg_apply(int Nm, int nc, const double* __restrict__ Ai, const double* __restrict__ cs,
           const float4* __restrict__ w, double* __restrict__ out, const SolveScalars* __restrict__ sc,
           int guard) {
    main_stream_priority();
    if (guard == 1 && sc->done) return;
    if (guard == 2 && !sc->done) return;
    const int b = blockIdx.x, l = threadIdx.x;
    double c[4] = {0.0, 0.0, 0.0, 0.0};
    for (int m = l; m < nc; m += kTileEdge) {
        const double sv = cs[m];
#pragma unroll
        for (int a = 0; a < 4; ++a) c[a] += Ai[(size_t)(4 * b + a) * nc + m] * sv;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) c[a] = wave_sum(c[a]);
    const int i = b * kTileEdge + l;
    if (i >= Nm) return;
    const float4 z = w[i];
    out[3 * Nm + i] += ((z.x * c[0] + z.y * c[1]) + z.z * c[2]) + z.w * c[3];
}

}  // namespace

void bj_free(capac_ctx* c) {
    void* p[] = {c->bj_inv, c->bj_r, c->cg_w, c->cg_act, c->cg_s, c->cg_A};
    for (void* q : p)
        if (q) (void)hipFree(q);
    c->bj_inv = NULL; c->bj_r = NULL; c->cg_w = NULL; c->cg_act = NULL; c->cg_s = NULL; c->cg_A = NULL;
    c->cg_n = 0;
    c->bj_valid = 0;
    c->bj_age = 0;
}

// the inverses and, up to kCgMax coarse unknowns, the coarse space's buffers
int bj_alloc_dev(capac_ctx* c) {
    bj_free(c);
    const int Nm = c->bc.Nm;
    if (c->bj_every <= 0 || Nm == 0) return 0;
    const int Nb = tile_blocks(Nm), nc = 4 * Nb;
    CHECK_HIP(hipMalloc(&c->bj_inv, (size_t)Nb * kBjUnk * kBjUnk * sizeof(float)));
    CHECK_HIP(hipMalloc(&c->bj_r, 3 * (size_t)Nm * sizeof(double)));
    if (CAPAC_BJ_COARSE && nc <= kCgMax) {
        CHECK_HIP(hipMalloc(&c->cg_w, (size_t)Nb * kTileEdge * sizeof(float4)));
        CHECK_HIP(hipMalloc(&c->cg_act, (size_t)nc * sizeof(int)));
        CHECK_HIP(hipMalloc(&c->cg_s, (size_t)nc * sizeof(double)));
        CHECK_HIP(hipMalloc(&c->cg_A, (size_t)nc * nc * sizeof(double)));
        c->cg_n = nc;
    }
    return 0;
}

size_t bj_bytes(int Nm) { return (size_t)tile_blocks(Nm) * kBjUnk * kBjUnk * sizeof(float); }

// inverts the diagonal blocks of the current FP32 relay (after cpim_assemble)
int bj_build(capac_ctx* c) {
    const int Nm = c->bc.Nm;
    if (Nm == 0 || !c->bj_inv) return 0;
    const int tiles = c->relay_layout == CAPAC_LAYOUT_TILES;
    const int Nb = tile_blocks(Nm);
    k_bj_build<<<Nb, kBjThreads, 0, c->stream>>>(c->relay32, tiles, Nb, c->relay_ld, c->r, c->metal_u, c->bc,
                                                 c->bj_inv, c->bj_r);
    if (c->cg_n) {
        float4* w = reinterpret_cast<float4*>(c->cg_w);
        k_cg_basis<<<Nb, kTileEdge, 0, c->stream>>>(c->r, c->bc, w, c->cg_act);
        k_cg_coarse<<<(unsigned)tile_count(Nb), 256, 0, c->stream>>>(c->relay32, tiles, Nb, c->relay_ld, w,
                                                                     c->cg_act, c->bc, c->cg_A, c->cg_n);
        k_cg_invert<<<1, kBjThreads, 0, c->stream>>>(c->cg_A, c->cg_n);
    }
    CHECK_HIP(hipGetLastError());
    return 0;
}

// block b's inverse as a row-major 256 x 256 FP32 matrix in host memory
int bj_copy_block(capac_ctx* c, int b, float* out) {
    if (b < 0 || b >= tile_blocks(c->bc.Nm)) {
        capac_set_error("block %d is outside the %d blocks of 64 metal atoms", b, tile_blocks(c->bc.Nm));
        return CAPAC_ERR_ARG;
    }
    if (!c->bj_inv || !c->bj_valid) {
        capac_set_error("no block inverses formed (block preconditioner off, or no assembly since it was set)");
        return CAPAC_ERR_ARG;
    }
    const size_t n = (size_t)kBjUnk * kBjUnk;
    float* tmp = (float*)malloc(n * sizeof(float));
    if (!tmp) { capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
    hipError_t e = hipMemcpyAsync(tmp, c->bj_inv + (size_t)b * n, n * sizeof(float), hipMemcpyDeviceToHost,
                                  c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        free(tmp);
        capac_set_error("block inverse copy: %s", hipGetErrorString(e));
        return CAPAC_ERR_HIP;
    }
    for (int rho = 0; rho < kBjUnk; ++rho)
        for (int sig = 0; sig < kBjUnk; ++sig)
            out[(size_t)rho * kBjUnk + sig] = tmp[(sig >> 2) * (4 * kBjUnk) + rho * 4 + (sig & 3)];
    free(tmp);
    return 0;
}

// the coarse space as last formed: its size, Ac^-1 (n x n, row-major) and
// the basis (float4 per block slot of 64 atoms), either output optional
int bj_copy_coarse(capac_ctx* c, int* n, double* ainv, float* basis) {
    *n = c->cg_n && c->bj_valid ? c->cg_n : 0;
    if (!*n) return 0;
    const int Nb = tile_blocks(c->bc.Nm);
    if (ainv)
        CHECK_HIP(hipMemcpyAsync(ainv, c->cg_A, (size_t)*n * *n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (basis)
        CHECK_HIP(hipMemcpyAsync(basis, c->cg_w, (size_t)Nb * kTileEdge * 4 * sizeof(float), hipMemcpyDeviceToHost,
                                 c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

void bj_apply(capac_ctx* c, const double* q, double* out, bool add, const SolveScalars* sc, int guard) {
    const int Nm = c->bc.Nm, Nb = tile_blocks(Nm);
    const int blocks = 4 * Nb + (c->n_mat > 4 * Nm ? 1 : 0);
    const float4* w = reinterpret_cast<const float4*>(c->cg_w);
    double* cs = c->cg_n ? c->cg_s : NULL;
    if (add)
        k_bj_apply<true><<<blocks, 256, 0, c->stream>>>(Nm, Nb, c->n_mat, c->bj_inv, q, out, w, cs, sc, guard);
    else
        k_bj_apply<false><<<blocks, 256, 0, c->stream>>>(Nm, Nb, c->n_mat, c->bj_inv, q, out, w, cs, sc, guard);
}

void cg_apply(capac_ctx* c, double* out, const SolveScalars* sc, int guard) {
    if (!c->cg_n) return;
    const int Nm = c->bc.Nm;
    k_cg_apply<<<tile_blocks(Nm), kTileEdge, 0, c->stream>>>(Nm, c->cg_n, c->cg_A, c->cg_s,
                                                             reinterpret_cast<const float4*>(c->cg_w), out, sc,
                                                             guard);
}
