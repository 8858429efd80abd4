// Jacobi-preconditioned BiCGSTAB on the relay matrix: the reference's
// mpi_precon_bicg_stab_COO (bicg_stab.c:253-573) with its exact control flow:
//   bTb < 1e-10 -> return, x unchanged                       (bicg_stab.c:300)
//   criteria = 1e-8 bTb, warm start from the resident x       (:303-306)
//   rho < 1e-10 at the first iteration -> return             (:368)
//   omega = (Kt, Ks)/(Kt, Kt)                                 (:483-488)
//   stop when rTr < criteria; after 1000 iterations accept rTr < 100 criteria
//   ("loosened"), else fail                                   (:539-571)
// Lagrange rows/columns of the neutrality constraint: the charge row of metal
// i gets + lambda_k(i); the row of nanoparticle k is sum_{i in k} q_i; their
// preconditioner entries are 1.
//
// Iteration n is four launches:
//   K1  v = A y                SpMV, per row-group partial (r0, v)
//   K2  alpha = rho/(r0, v)    multi-block vector kernel: s = r - alpha v,
//       z = K s, partial (r0, s); it also finishes the stop test of
//       iteration n-1 from the partials (r, r) of K4
//   K3  t = A z                SpMV, partials (Kt, z), (Kt, Kt), (r0, t)
//   K4  omega = (Kt,z)/(Kt,Kt) multi-block: x += alpha y + omega z,
//       r = s - omega t, partial (r, r); rho' = (r0, r) taken as
//       (r0, s) - omega (r0, t) so that beta and p = r + beta (p - omega v),
//       y = K p follow in the same pass
// Every block of a vector kernel reduces the partials of the previous launch
// itself, in the same fixed order, so all blocks hold bitwise-equal scalars
// without a grid barrier (on this GPU the eight XCDs have separate L2
// caches, and an in-kernel device-scope fence costs more than a launch).
// The stop test of iteration n runs one launch later, in K2 of n+1, so one
// speculative SpMV is spent at the end of a solve.  All scalars live on the
// device; a `done` flag turns the remaining launches of a batch into no-ops
// and the host synchronises once per batch.
//
// Relay formats:
//   RelayM64  FP64, no stored matrix: each block is formed from the
//             positions inside the product (the reference's entries)
//   RelayG32  four FP32 coefficients per ordered pair plus the separation
//             formed from the FP64 positions; products accumulate in FP64
//             with FP64 vectors; diagonal blocks applied analytically
//   tiles     the same coefficients, each unordered pair once (relay_tile.h)
// The FP64 path runs the reference's BiCGSTAB on RelayM64.  The FP32-refined
// path runs GMRES on its FP32 relay (section below), then checks the true
// residual b - A x with the FP64 relay recomputed on the fly and, while the
// reference's criterion fails on it, solves for the correction on the FP32
// relay and adds it (iterative refinement).
#include "capac_internal.h"
#include "relay.h"
#include "precond.h"
#include "relay_tile.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>

// GMRES restart length: the columns kept before the solve restarts from its
// updated x (Au2057 with the quadratic warm start needs ~20-30)
constexpr int kGmM = 40;

// GMRES state on the device (kernels below)
struct GmState {
    double nrm[kGmM + 1];           // |q_i|; the basis vector v_i = q_i / nrm[i]
    double H[kGmM][kGmM + 1];       // column j: h_0j .. h_jj as formed, rotated in place into R
    double cs[kGmM], sn[kGmM];      // Givens rotations
    double g[kGmM + 1];             // rotated right-hand side |r0| e_0
    double y[kGmM];                 // the update's coefficients once a solve or restart is due
    int ny;                         // columns in y (0: no update pending)
    int pad;
};

struct Solver {
    int n_mat, Nm, n_np, n_groups, n_vb;
    double *r0, *r, *p, *v, *s, *t, *y, *z, *K, *Ax;
    double* part1;          // n_groups: (r0, v)
    double* part3;          // 3 n_groups: (Kt, z), (Kt, Kt), (r0, t)
    double* part2;          // n_vb: (r0, s)
    double* part4;          // n_vb: (r, r)
    double *d, *rr, *part_res;   // refinement: correction, FP64 residual, its partials
    // GMRES (the FP32-refined path's inner solver): the basis, the per-block
    // partials of its dot products, the Hessenberg column state
    double* Q;              // (kGmM + 1) n_mat: unnormalised basis vectors q_i
    double* part_h;         // ph_rows (kGmM + 1)
    int ph_rows;            // rows of part_h: n_vb, or the fused column's Nb + n_np when more
    GmState* gm;            // device
    const float* bj;        // block-Jacobi inverses (precond.hip), or NULL: the Jacobi K
    // inside the context's readback buffer, so that the synchronisation at
    // the end of a batch also brings back the neighbour-row flags
    double* res;            // device: [0] (rr, rr)
    SolveScalars* sc;       // device
    SolveScalars* sc_host;  // pinned mirror
    double* res_host;       // pinned mirror
    bool in_red;            // sc and res inside the context's readback buffer
};

namespace {

constexpr int kVecThreads = 256;       // vector kernels: one row per thread

// FP64 blocks formed from the positions in the product (VALU work, one
// row-group per block: R = 1); acc skips j == i (diag applies that block)
struct RelayM64 {
    const double* r;
    BoxConst bc;
    template <int R>
    struct Rows {
        const double* r;
        const BoxConst* bc;
        int i, N;
        double xi, yi, zi;
        template <int U>
        __device__ __forceinline__ void acc_n(int j, int Nm, const double* __restrict__ x, double (&o)[R][4]) const {
            static_assert(U == 1, "one column per lane and pass");
            acc(j, x[3 * j], x[3 * j + 1], x[3 * j + 2], x[3 * Nm + j], o);
        }
        template <int U>
        __device__ __forceinline__ void acc_compact(int j, int Nm, const double* __restrict__ x,
                                                    double (&o)[R][4]) const {
            acc_n<U>(j, Nm, x, o);
        }
        __device__ __forceinline__ void acc(int j, double x0, double x1, double x2, double x3,
                                            double (&o)[R][4]) const {
            static_assert(R == 1, "the FP64 product takes one row-group per block");
            if (j == i) return;
            double blk[16];
            relay_block64_sep(r[j] - xi, r[N + j] - yi, r[2 * N + j] - zi, *bc, blk);
            o[0][0] += blk[0] * x0 + blk[1] * x1 + blk[2] * x2 + blk[3] * x3;
            o[0][1] += blk[4] * x0 + blk[5] * x1 + blk[6] * x2 + blk[7] * x3;
            o[0][2] += blk[8] * x0 + blk[9] * x1 + blk[10] * x2 + blk[11] * x3;
            o[0][3] += blk[12] * x0 + blk[13] * x1 + blk[14] * x2 + blk[15] * x3;
        }
    };
    template <int R>
    __device__ __forceinline__ Rows<R> rows(int i0) const {
        const int N = bc.N;
        return Rows<R>{r, &bc, i0, N, r[i0], r[N + i0], r[2 * N + i0]};
    }
    static constexpr int kUnroll = 1;      // columns per lane and pass
    __device__ __forceinline__ bool compact() const { return false; }
    __device__ __forceinline__ void diag(double x0, double x1, double x2, double x3, double o[4]) const {
        o[0] += bc.inv_polar * x0;
        o[1] += bc.inv_polar * x1;
        o[2] += bc.inv_polar * x2;
        o[3] += -bc.inv_capac * x3;
    }
};

#ifndef CAPAC_G32_UNROLL
#define CAPAC_G32_UNROLL 2
#endif

// FP32 coefficients g_a, g_b, c3, cq per ordered pair (relay.h) for R
// consecutive row-groups at once: column j's vector entries and position are
// loaded once for all R rows, each row adds its block from its own four
// coefficients and the separation formed from the FP64 positions.  The
// relay holds zero coefficient rows up to a multiple of kRelay32RowPad (R
// must not exceed it), so
// a block's last rows need no test (their position index is clamped to a
// metal atom, and zero coefficients add exactly zero); the diagonal pair's
// coefficients are zero too, so j == i needs no test either.
struct RelayG32 {
    const float* a;
    int ld;
    const double* r;
    const double* u;                // metal separations from metal atom 0 (k_metal_unwrap)
    BoxConst bc;
    template <int R>
    struct Rows {
        const float* cp;            // plane 0 of row-group i0
        int ld, N;
        const double* r;
        const double* u;
        const BoxConst* bc;
        double xi[R], yi[R], zi[R];
        double ux[R], uy[R], uz[R];
        // U columns j, j + 256, ... per lane: every load of the U columns is
        // issued before the first block is applied, so a wavefront keeps U
        // columns' worth of loads in flight.  A column past the end is read
        // at Nm - 1 with zero vector entries (its blocks add exactly zero).
        template <int U>
        __device__ __forceinline__ void acc_n(int j0, int Nm, const double* __restrict__ x,
                                              double (&o)[R][4]) const {
            float c[U][R][kRelay32Planes];
            double xv[U][4], pj[U][3];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = j0 + 256 * u;
                const bool in = j < Nm;
                const int jj = in ? j : Nm - 1;
#pragma unroll
                for (int q = 0; q < R; ++q)
#pragma unroll
                    for (int k = 0; k < kRelay32Planes; ++k)
                        c[u][q][k] = cp[(size_t)(q * kRelay32Planes + k) * ld + jj];
                xv[u][0] = x[3 * jj]; xv[u][1] = x[3 * jj + 1]; xv[u][2] = x[3 * jj + 2];
                xv[u][3] = x[3 * Nm + jj];
                if (!in) xv[u][0] = xv[u][1] = xv[u][2] = xv[u][3] = 0.0;
                pj[u][0] = r[jj]; pj[u][1] = r[N + jj]; pj[u][2] = r[2 * N + jj];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    double dx = pj[u][0] - xi[q], dy = pj[u][1] - yi[q], dz = pj[u][2] - zi[q];
                    pbc12(*bc, dx, dy, dz);
                    dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
                    block32_apply(dx, dy, dz, c[u][q][0], c[u][q][1], c[u][q][2], c[u][q][3], xv[u][0], xv[u][1],
                                  xv[u][2], xv[u][3], o[q]);
                }
        }
        // the same when the metal is compact: the separation is d_j - d_i
        // from the stored separations in bohr, no image test per pair
        template <int U>
        __device__ __forceinline__ void acc_compact(int j0, int Nm, const double* __restrict__ x,
                                                    double (&o)[R][4]) const {
            float c[U][R][kRelay32Planes];
            double xv[U][4], pj[U][3];
#pragma unroll
            for (int w = 0; w < U; ++w) {
                const int j = j0 + 256 * w;
                const bool in = j < Nm;
                const int jj = in ? j : Nm - 1;
#pragma unroll
                for (int q = 0; q < R; ++q)
#pragma unroll
                    for (int k = 0; k < kRelay32Planes; ++k)
                        c[w][q][k] = cp[(size_t)(q * kRelay32Planes + k) * ld + jj];
                xv[w][0] = x[3 * jj]; xv[w][1] = x[3 * jj + 1]; xv[w][2] = x[3 * jj + 2];
                xv[w][3] = x[3 * Nm + jj];
                if (!in) xv[w][0] = xv[w][1] = xv[w][2] = xv[w][3] = 0.0;
                const double4 d = reinterpret_cast<const double4*>(u)[jj];
                pj[w][0] = d.x; pj[w][1] = d.y; pj[w][2] = d.z;
            }
#pragma unroll
            for (int w = 0; w < U; ++w)
#pragma unroll
                for (int q = 0; q < R; ++q)
                    block32_apply(pj[w][0] - ux[q], pj[w][1] - uy[q], pj[w][2] - uz[q], c[w][q][0], c[w][q][1],
                                  c[w][q][2], c[w][q][3], xv[w][0], xv[w][1], xv[w][2], xv[w][3], o[q]);
        }
    };
    template <int R>
    __device__ __forceinline__ Rows<R> rows(int i0) const {
        static_assert(R <= kRelay32RowPad, "row block larger than the relay's row padding");
        const int N = bc.N, Nm = bc.Nm;
        Rows<R> w;
        w.cp = a + (size_t)i0 * kRelay32Planes * ld;
        w.ld = ld; w.N = N; w.r = r; w.u = u; w.bc = &bc;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = i0 + q < Nm ? i0 + q : Nm - 1;
            w.xi[q] = r[i]; w.yi[q] = r[N + i]; w.zi[q] = r[2 * N + i];
            w.ux[q] = u[4 * i]; w.uy[q] = u[4 * i + 1]; w.uz[q] = u[4 * i + 2];
        }
        return w;
    }
    __device__ __forceinline__ bool compact() const { return u[4 * bc.Nm] != 0.0; }
    __device__ __forceinline__ void diag(double x0, double x1, double x2, double x3, double o[4]) const {
        o[0] += bc.inv_polar * x0;
        o[1] += bc.inv_polar * x1;
        o[2] += bc.inv_polar * x2;
        o[3] -= bc.inv_capac * x3;
    }
    static constexpr int kUnroll = CAPAC_G32_UNROLL;
};

// fixed-order block sum (blockDim a multiple of 64, <= 1024)
__device__ double block_sum(double v, double* s_w) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) s_w[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += s_w[w];
    return t;
}

// fixed-order sum of `n` strided values src[k*stride + off] over one block
__device__ double block_reduce(const double* __restrict__ src, int n, int stride, int off, double* s_w) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) acc += src[(size_t)k * stride + off];
    return block_sum(acc, s_w);
}

// NV fixed-order block sums at once (one LDS round); s_w holds 16 x NV
template <int NV>
__device__ void block_sum_n(double v[NV], double* s_w) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = wave_sum(v[c]);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NV; ++c) s_w[wave * NV + c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += s_w[w * NV + c];
        v[c] = t;
    }
}

// the outputs of row-group g from its four row sums o (see k_spmv)
template <class M>
__device__ __forceinline__ void spmv_group_out(const M& m, int g, int Nm, const int* __restrict__ np_of,
                                               const double* __restrict__ x, double* __restrict__ y,
                                               int mode, const double* __restrict__ r0,
                                               const double* __restrict__ K, double* __restrict__ part,
                                               double o[4]) {
    if (g >= Nm) {
        const int row = 4 * Nm + (g - Nm);
        const double val = o[3];
        y[row] = val;
        if (mode == 1) part[g] = r0[row] * val;
        if (mode == 2) {
            double kt = K[row] * val;
            part[3 * g] = kt * x[row];
            part[3 * g + 1] = kt * kt;
            part[3 * g + 2] = r0[row] * val;
        }
        return;
    }
    const int i = g;
    m.diag(x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * Nm + i], o);
    const int np = np_of[i];
    if (np >= 0) o[3] += x[4 * Nm + np];
    const int rows[4] = {3 * i, 3 * i + 1, 3 * i + 2, 3 * Nm + i};
#pragma unroll
    for (int c = 0; c < 4; ++c) y[rows[c]] = o[c];
    if (mode == 1) {
        double d = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) d += r0[rows[c]] * o[c];
        part[g] = d;
    } else if (mode == 2) {
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double kt = K[rows[c]] * o[c];
            d0 += kt * x[rows[c]];
            d1 += kt * kt;
            d2 += r0[rows[c]] * o[c];
        }
        part[3 * g] = d0;
        part[3 * g + 1] = d1;
        part[3 * g + 2] = d2;
    }
}

// y = A x over row-groups: group g < Nm is metal i's four rows (3 dipole, 1
// charge), group Nm + k the Lagrange row of nanoparticle k.
//   mode 0: y only
//   mode 1: part[g] = (r0, y)_g
//   mode 2: part[3g..3g+2] = (K y, x)_g, (K y, K y)_g, (r0, y)_g
// One workgroup of 256 threads per R consecutive metal row-groups (the
// threads stride over the columns j), then one per Lagrange row; the row
// sums are fixed-order wavefront and block reductions.
template <class M, int R>
__global__ void __launch_bounds__(256)
k_spmv(M m, int Nm, int n_np, const int* __restrict__ np_of, const int* __restrict__ np_first,
       const int* __restrict__ np_last, const double* __restrict__ x, double* __restrict__ y, int mode,
       const double* __restrict__ r0, const double* __restrict__ K, double* __restrict__ part,
       const SolveScalars* __restrict__ sc) {
    main_stream_priority();
    __shared__ double s_w[4][4 * R];
    if (sc && sc->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n_mblk = (Nm + R - 1) / R;
    double o[R][4];
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) o[q][c] = 0.0;
    const bool lagrange = (int)blockIdx.x >= n_mblk;
    const int i0 = blockIdx.x * R;
    if (lagrange) {
        const int k = blockIdx.x - n_mblk;
        for (int mi = np_first[k] + threadIdx.x; mi <= np_last[k]; mi += 256) o[0][3] += x[3 * Nm + mi];
    } else {
        const auto rows = m.template rows<R>(i0);
        constexpr int U = M::kUnroll;
        if (m.compact())
            for (int j = threadIdx.x; j < Nm; j += 256 * U) rows.template acc_compact<U>(j, Nm, x, o);
        else
            for (int j = threadIdx.x; j < Nm; j += 256 * U) rows.template acc_n<U>(j, Nm, x, o);
    }
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double v = wave_sum(o[q][c]);
            if (lane == 0) s_w[wave][4 * q + c] = v;
        }
    __syncthreads();
    const int q = threadIdx.x;
    if (lagrange) {
        if (q != 0) return;
    } else if (q >= R || i0 + q >= Nm) {
        return;
    }
    double oq[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) oq[c] = s_w[0][4 * q + c] + s_w[1][4 * q + c] + s_w[2][4 * q + c] + s_w[3][4 * q + c];
    const int g = lagrange ? Nm + (blockIdx.x - n_mblk) : i0 + q;
    spmv_group_out(m, g, Nm, np_of, x, y, mode, r0, K, part, oq);
}

// Row block b's outputs from its Nb partials P[b][q] (relay_tile.h):
// thread (g, l) sums the four components of row l over the source blocks
// q = g, g + G, ... (G = blockDim / 64 groups), kSumBatch blocks' loads
// issued before their first add, so the partials arrive in one memory round
// trip per G kSumBatch source blocks; the G groups are added in order, then
// row l's diagonal block,
// Lagrange column and outputs follow as k_spmv writes them.  Every thread of
// the workgroup calls it (one barrier inside); s_o holds G x 4 x 64 doubles.
constexpr int kSumBatch = 8;            // source blocks per thread and round trip
__device__ void tile_block_out(int b, int Nb, int Nm, const int* __restrict__ np_of,
                               const double* __restrict__ P, const BoxConst& bc, const double* __restrict__ x,
                               double* __restrict__ y, int mode, const double* __restrict__ r0,
                               const double* __restrict__ K, double* __restrict__ part,
                               double (*s_o)[4][kTileEdge]) {
    const int G = blockDim.x >> 6, g = threadIdx.x >> 6, l = threadIdx.x & 63;
    const double* src = P + (size_t)b * Nb * (4 * kTileEdge) + l;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q0 = g; q0 < Nb; q0 += G * kSumBatch) {
        double v[kSumBatch][4];
#pragma unroll
        for (int t = 0; t < kSumBatch; ++t) {
            const int q = q0 + G * t;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                v[t][c] = q < Nb ? src[(size_t)q * (4 * kTileEdge) + c * kTileEdge] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < kSumBatch; ++t)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] += v[t][c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) s_o[g][c][l] = acc[c];
    __syncthreads();
    const int i = b * kTileEdge + threadIdx.x;
    if (threadIdx.x >= kTileEdge || i >= Nm) return;
    double o[4] = {0.0, 0.0, 0.0, 0.0};
    for (int h = 0; h < G; ++h)
#pragma unroll
        for (int c = 0; c < 4; ++c) o[c] += s_o[h][c][threadIdx.x];
    RelayG32 m{NULL, 0, NULL, NULL, bc};
    spmv_group_out(m, i, Nm, np_of, x, y, mode, r0, K, part, o);
}

// y = A x on the TILES layout (relay_tile.h), first of two launches.  One
// workgroup per tile (I, J): the lanes are the rows i of block I, and at
// diagonal k lane l meets column j = J*64 + ((l + k) & 63), so at every k the
// wavefront covers 64 distinct pairs.  Each pair is applied to row i (from
// x_j) and to row j (from x_i, the block transposed: r -> -r flips the
// charge-dipole terms).  The row sums stay in the lane; the column sums go
// to the wavefront's own column array in LDS through FP64 LDS atomic adds
// that return nothing (the lane does not wait for them): at one k the 64
// lanes add to 64 distinct columns, and one wavefront's LDS operations
// complete in program order, so every column's sum is formed in a fixed
// order.  The kTileWaves wavefronts take consecutive runs of diagonals.
// Their sums are added in a fixed order and written as partials
// P[dst][src][c][l] (64 rows of block dst, component c, from block src):
// P[I][J] the row sums, P[J][I] the column sums, P[I][I] the row sums of a
// diagonal tile.  Workgroups past the tiles form the Lagrange rows.
// Wavefronts per tile: 4 (16 diagonals each, at most 168 VGPRs) lets three
// workgroups share a compute unit, so the 561 tiles of Au2057 all start at
// once.  Measured at Au2057 (product alone / MD step): 4 waves 15.7 us /
// 1.91 ms; 8 waves (at most 128 VGPRs) 16.1 us / 2.21 ms, their larger
// workgroups finding fewer free slots beside the pair kernel.
#ifndef CAPAC_TILE_WAVES
#define CAPAC_TILE_WAVES 4
#endif
constexpr int kTileWaves = CAPAC_TILE_WAVES;            // wavefronts per tile
constexpr int kTileThreads = kTileWaves * CAPAC_WAVE;
constexpr int kTileSteps = kTileEdge / kTileWaves;      // diagonals per wavefront

// the wavefront's diagonals k0 .. k0 + kTileSteps - 1, their coefficients cf
// already loaded.  Column j's entries for the next diagonal are read before
// this diagonal's column adds are issued: the adds then never stand between
// a read and its use (LDS operations complete in order).
template <bool Compact>
__device__ __forceinline__ void tile_apply(const float4 (&cf)[kTileSteps], int k0, const BoxConst& bc,
                                           const double (*s_x)[kTileEdge], const double (*s_p)[kTileEdge],
                                           const double xi[4], const double pi[3], int lane, double orow[4],
                                           double (*s_col)[kTileEdge]) {
    int jl = (lane + k0) & (kTileEdge - 1);
    double xj[4] = {s_x[0][jl], s_x[1][jl], s_x[2][jl], s_x[3][jl]};
    double pj[3] = {s_p[0][jl], s_p[1][jl], s_p[2][jl]};
#pragma unroll
    for (int s = 0; s < kTileSteps; ++s) {
        double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
        if (!Compact) {
            // pbc_12 with both candidates formed and selected (same results;
            // its branches inside this unrolled loop double the kernel's
            // registers)
            const double ax = dx + bc.box[0], bx = dx - bc.box[0];
            const double ay = dy + bc.box[1], by = dy - bc.box[1];
            const double az = dz + bc.box[2], bz = dz - bc.box[2];
            dx = dx < -bc.half[0] ? ax : (dx > bc.half[0] ? bx : dx);
            dy = dy < -bc.half[1] ? ay : (dy > bc.half[1] ? by : dy);
            dz = dz < -bc.half[2] ? az : (dz > bc.half[2] ? bz : dz);
            dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
        }
        const double ga = cf[s].x, gb = cf[s].y, c3 = cf[s].z, cq = cf[s].w;
        // row i from column j (block (i, j) of relay.h)
        const double rpj = dx * xj[0] + dy * xj[1] + dz * xj[2];
        const double gr = ga * rpj + c3 * xj[3];
        orow[0] += gr * dx + gb * xj[0];
        orow[1] += gr * dy + gb * xj[1];
        orow[2] += gr * dz + gb * xj[2];
        orow[3] += cq * xj[3] - c3 * rpj;
        // row j from column i: block (j, i) has separation -r
        const double rpi = dx * xi[0] + dy * xi[1] + dz * xi[2];
        const double gc = ga * rpi - c3 * xi[3];
        const double oc[4] = {gc * dx + gb * xi[0], gc * dy + gb * xi[1], gc * dz + gb * xi[2],
                              cq * xi[3] + c3 * rpi};
        const int jc = jl;
        if (s + 1 < kTileSteps) {
            jl = (jl + 1) & (kTileEdge - 1);
#pragma unroll
            for (int c = 0; c < 4; ++c) xj[c] = s_x[c][jl];
#pragma unroll
            for (int c = 0; c < 3; ++c) pj[c] = s_p[c][jl];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            (void)__hip_atomic_fetch_add(&s_col[c][jc], oc[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// The coarse correction the tiled product adds to its input (GMRES with the
// block preconditioner): x's charges take w(i) . c_b, c = Ac^-1 s, s = Z^T q
// as the update left it; the corrected vector goes to x2 for the sums launch.
// nc = 0: none.  With one, the sums launch forms the Lagrange rows from x2.
struct CoarseArgs {
    const double* A;        // Ac^-1, nc x nc
    const double* s;        // Z^T q, nc
    const float4* w;        // basis entries per metal atom
    double* x2;             // the corrected input (n_mat)
    int nc;
};

#ifndef CAPAC_TILE_MINW
#define CAPAC_TILE_MINW 3
#endif
// CAPAC_TILE_FUSED=1 forms the row blocks' outputs inside the tile launch:
// the workgroup that delivers a block's last partial sums it, after a
// device-scope release fence, a counter add and an acquire fence.  It saves
// the second launch but measured 68 us per product at Au2057 against
// 16.3 us for the two launches (every workgroup's release fence writes back
// its XCD's L2), so it is off; it stays buildable as that measurement.
#ifndef CAPAC_TILE_FUSED
#define CAPAC_TILE_FUSED 0
#endif
__global__ void __launch_bounds__(kTileThreads, CAPAC_TILE_MINW)
k_spmv_tiles(const float4* __restrict__ tiles, int Nb, int n_tiles, int Nm, int n_np,
             const int* __restrict__ np_of, const int* __restrict__ np_first, const int* __restrict__ np_last,
             const double* __restrict__ r, const double* __restrict__ u, BoxConst bc,
             const double* __restrict__ x, double* __restrict__ y, double* __restrict__ P, int* __restrict__ cnt,
             int mode, const double* __restrict__ r0, const double* __restrict__ K, double* __restrict__ part,
             const SolveScalars* __restrict__ sc, CoarseArgs cg) {
    main_stream_priority();
    // component-major, so that the lanes' reads at consecutive rows hit
    // consecutive banks
    __shared__ double s_x[2][4][kTileEdge];
    __shared__ double s_p[2][3][kTileEdge];
    __shared__ double s_acc[2][kTileWaves][4][kTileEdge];    // [row/column][wave][component][lane]
    __shared__ double s_cc[2][4];
    if (sc && sc->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if ((int)blockIdx.x >= n_tiles) {
        // Lagrange row of nanoparticle k: the sum of its metal charges (with
        // a coarse correction the sums launch forms these rows instead)
        const int k = blockIdx.x - n_tiles;
        double o[4] = {0.0, 0.0, 0.0, 0.0};
        for (int mi = np_first[k] + threadIdx.x; mi <= np_last[k]; mi += kTileThreads) o[3] += x[3 * Nm + mi];
        double v = wave_sum(o[3]);
        if (lane == 0) s_acc[0][0][0][wave] = v;
        __syncthreads();
        if (threadIdx.x != 0) return;
        o[3] = 0.0;
        for (int w = 0; w < kTileWaves; ++w) o[3] += s_acc[0][0][0][w];
        RelayG32 m{NULL, 0, r, u, bc};
        spmv_group_out(m, Nm + k, Nm, np_of, x, y, mode, r0, K, part, o);
        return;
    }
    int I, J;
    tile_coords(blockIdx.x, Nb, I, J);
    // the wavefront's coefficients first: their HBM stream then overlaps the
    // staging of the vectors below
    const int k0 = kTileSteps * wave;
    float4 cf[kTileSteps];
    {
        const float4* cp = tiles + (size_t)blockIdx.x * kTileSlots + (size_t)k0 * kTileEdge + lane;
#pragma unroll
        for (int s = 0; s < kTileSteps; ++s) cf[s] = cp[s * kTileEdge];
    }
    // with a coarse correction, waves 2 and 3 form the coarse values of
    // blocks I and J (rows 4I.. and 4J.. of Ac^-1 times Z^T q) while waves 0
    // and 1 stage the vectors; the staged charges then take w(i) . c
    float4 zw = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cg.nc) {
        if (threadIdx.x >= 2 * kTileEdge) {
            const int side = wave - 2, blk = side ? J : I;
            double cc[4] = {0.0, 0.0, 0.0, 0.0};
            // every load of the row segments issued before the first add
            double av[kCgMax / kTileEdge][4], sv[kCgMax / kTileEdge];
#pragma unroll
            for (int it = 0; it < kCgMax / kTileEdge; ++it) {
                const int m = lane + kTileEdge * it;
                sv[it] = m < cg.nc ? cg.s[m] : 0.0;
#pragma unroll
                for (int a = 0; a < 4; ++a) av[it][a] = m < cg.nc ? cg.A[(size_t)(4 * blk + a) * cg.nc + m] : 0.0;
            }
#pragma unroll
            for (int it = 0; it < kCgMax / kTileEdge; ++it)
#pragma unroll
                for (int a = 0; a < 4; ++a) cc[a] += av[it][a] * sv[it];
#pragma unroll
            for (int a = 0; a < 4; ++a) cc[a] = wave_sum(cc[a]);
            if (lane == 0)
#pragma unroll
                for (int a = 0; a < 4; ++a) s_cc[side][a] = cc[a];
        } else {
            const int m = ((threadIdx.x >> 6) ? J : I) * kTileEdge + lane;
            if (m < Nm) zw = cg.w[m];
        }
    }
    // the compact flag and both position forms are loaded together, so the
    // staging costs one memory round trip
    const double flag = u[4 * Nm];
    if (threadIdx.x < 2 * kTileEdge) {
        // x and positions of blocks I (side 0) and J (side 1); zeros past Nm
        const int side = threadIdx.x >> 6;
        const int m = (side ? J : I) * kTileEdge + lane;
        double xv[4] = {0.0, 0.0, 0.0, 0.0}, pv[3] = {0.0, 0.0, 0.0};
        if (m < Nm) {
            xv[0] = x[3 * m]; xv[1] = x[3 * m + 1]; xv[2] = x[3 * m + 2]; xv[3] = x[3 * Nm + m];
            const double4 uv = reinterpret_cast<const double4*>(u)[m];
            const double rx = r[m], ry = r[bc.N + m], rz = r[2 * bc.N + m];
            const bool cmp = flag != 0.0;
            pv[0] = cmp ? uv.x : rx; pv[1] = cmp ? uv.y : ry; pv[2] = cmp ? uv.z : rz;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) s_x[side][c][lane] = xv[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) s_p[side][c][lane] = pv[c];
    }
    __syncthreads();
    const bool diag = I == J, compact = flag != 0.0;
    if (cg.nc) {
        if (threadIdx.x < 2 * kTileEdge) {
            const int side = threadIdx.x >> 6;
            const double* cc = s_cc[side];
            s_x[side][3][lane] += ((zw.x * cc[0] + zw.y * cc[1]) + zw.z * cc[2]) + zw.w * cc[3];
            // the diagonal tile keeps its block's corrected vector for the sums
            // launch; tile 0 also copies the Lagrange unknowns there
            const int m = I * kTileEdge + lane;
            if (diag && side == 0 && m < Nm) {
                cg.x2[3 * m] = s_x[0][0][lane];
                cg.x2[3 * m + 1] = s_x[0][1][lane];
                cg.x2[3 * m + 2] = s_x[0][2][lane];
                cg.x2[3 * Nm + m] = s_x[0][3][lane];
            }
            if (blockIdx.x == 0 && side == 1)
                for (int k = lane; k < n_np; k += kTileEdge) cg.x2[4 * Nm + k] = x[4 * Nm + k];
        }
        __syncthreads();
    }
    const double xi[4] = {s_x[0][0][lane], s_x[0][1][lane], s_x[0][2][lane], s_x[0][3][lane]};
    const double pi[3] = {s_p[0][0][lane], s_p[0][1][lane], s_p[0][2][lane]};
    double orow[4] = {0.0, 0.0, 0.0, 0.0};
    // this wavefront's column sums (its LDS operations run in order, so the
    // zeros land before its first add)
#pragma unroll
    for (int c = 0; c < 4; ++c) s_acc[1][wave][c][lane] = 0.0;
    if (compact) tile_apply<true>(cf, k0, bc, s_x[1], s_p[1], xi, pi, lane, orow, s_acc[1][wave]);
    else tile_apply<false>(cf, k0, bc, s_x[1], s_p[1], xi, pi, lane, orow, s_acc[1][wave]);
#pragma unroll
    for (int c = 0; c < 4; ++c) s_acc[0][wave][c][lane] = orow[c];
    __syncthreads();
    if (threadIdx.x < 4 * kTileEdge) {
        const int c = threadIdx.x >> 6;            // component; lane = row in the block
        double rs = 0.0, cs = 0.0;
#pragma unroll
        for (int w = 0; w < kTileWaves; ++w) {
            rs += s_acc[0][w][c][lane];
            cs += s_acc[1][w][c][lane];
        }
        const size_t blk = 4 * kTileEdge;
        if (diag) {
            // every ordered pair of the block is a slot of its own here: the
            // row sums are complete and the column sums (formed alongside, as
            // in the other tiles) are not used
            P[((size_t)I * Nb + I) * blk + c * kTileEdge + lane] = rs;
        } else {
            P[((size_t)I * Nb + J) * blk + c * kTileEdge + lane] = rs;
            P[((size_t)J * Nb + I) * blk + c * kTileEdge + lane] = cs;
        }
    }
#if CAPAC_TILE_FUSED
    // The workgroup that delivers a row block's last partial forms that
    // block's outputs (each block receives exactly Nb partials per launch):
    // release fence, one counter add per block written, and the last one
    // in acquires the others' partials.  The sums are taken in the same
    // fixed order whichever workgroup takes them; the counter is cleared for
    // the next launch.
    __shared__ int s_last[2];
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        s_last[0] = atomicAdd(&cnt[I], 1) == Nb - 1;
        s_last[1] = !diag && atomicAdd(&cnt[J], 1) == Nb - 1;
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        if (!s_last[side]) continue;
        const int b = side ? J : I;
        __threadfence();
        // s_acc is free again: its first 8 KB hold the block sums
        tile_block_out(b, Nb, Nm, np_of, P, bc, x, y, mode, r0, K, part,
                       reinterpret_cast<double (*)[4][kTileEdge]>(&s_acc[0][0][0][0]));
        if (threadIdx.x == 0) cnt[b] = 0;
        __syncthreads();
    }
#endif
}

// Second launch (without CAPAC_TILE_FUSED): one workgroup per block of 64
// metal atoms forms its rows' outputs from the partials.
// 256 threads (four groups of source blocks): in MD this launch runs beside
// the pair kernel, and a larger workgroup can find no compute unit with room
// for it until the pair kernel drains (a 1024-thread version waited 775 us
// behind it at Au2057)
constexpr int kSumThreads = 256;
__global__ void __launch_bounds__(kSumThreads)
k_spmv_tsum(int Nb, int Nm, const int* __restrict__ np_of, const double* __restrict__ P, BoxConst bc,
            const double* __restrict__ x, double* __restrict__ y, int mode, const double* __restrict__ r0,
            const double* __restrict__ K, double* __restrict__ part, const SolveScalars* __restrict__ sc,
            const int* __restrict__ np_first, const int* __restrict__ np_last) {
    main_stream_priority();
    __shared__ double s_o[kSumThreads / 64][4][kTileEdge];        // [group][component][row]
    if (sc && sc->done) return;
    if ((int)blockIdx.x >= Nb) {
        // Lagrange row of nanoparticle k (when the tile launch leaves them here)
        const int k = blockIdx.x - Nb, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        double o[4] = {0.0, 0.0, 0.0, 0.0};
        for (int mi = np_first[k] + threadIdx.x; mi <= np_last[k]; mi += kSumThreads) o[3] += x[3 * Nm + mi];
        const double v = wave_sum(o[3]);
        if (lane == 0) s_o[0][0][wave] = v;
        __syncthreads();
        if (threadIdx.x != 0) return;
        o[3] = 0.0;
        for (int w = 0; w < kSumThreads / 64; ++w) o[3] += s_o[0][0][w];
        RelayG32 m{NULL, 0, NULL, NULL, bc};
        spmv_group_out(m, Nm + k, Nm, np_of, x, y, mode, r0, K, part, o);
        return;
    }
    tile_block_out(blockIdx.x, Nb, Nm, np_of, P, bc, x, y, mode, r0, K, part, s_o);
}

// start of BiCGSTAB (bicg_stab.c:294-404) in two launches.  k_bicg_init,
// one thread per row: r0 = r = b - Ax, v = 0, K = 1/diag, the first
// direction p = r + rho (0 - 0) = r and y = K p, and per-block partial sums
// of (b, b) and (r, r).  k_bicg_scalars, one workgroup: bTb, criteria, rho =
// (r0, r) and the two early exits, which mark the solve done before any
// iteration runs (x unchanged).
__global__ void __launch_bounds__(kVecThreads)
k_bicg_init(int n_mat, int n4, const double* __restrict__ b, const double* __restrict__ Ax,
            const double* __restrict__ diag, double* __restrict__ r0, double* __restrict__ r,
            double* __restrict__ p, double* __restrict__ v, double* __restrict__ K,
            double* __restrict__ y, double* __restrict__ part_bb, double* __restrict__ part_rr) {
    main_stream_priority();
    __shared__ double s_w[16 * 2];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (k < n_mat) {
        const double bk = b[k];
        const double rk = bk - Ax[k];
        const double Kk = k < n4 ? 1.0 / diag[k] : 1.0;
        r0[k] = rk;
        r[k] = rk;
        v[k] = 0.0;
        K[k] = Kk;
        p[k] = rk;
        y[k] = rk * Kk;
        acc[0] = bk * bk;
        acc[1] = rk * rk;
    }
    block_sum_n<2>(acc, s_w);
    if (threadIdx.x == 0) {
        part_bb[blockIdx.x] = acc[0];
        part_rr[blockIdx.x] = acc[1];
    }
}

__global__ void __launch_bounds__(256)
k_bicg_scalars(int n_vb, const double* __restrict__ part_bb, const double* __restrict__ part_rr,
               SolveScalars* sc, double crit_factor) {
    main_stream_priority();
    __shared__ double s_w[16 * 2];
    double red[2] = {0.0, 0.0};
    for (int k = threadIdx.x; k < n_vb; k += blockDim.x) {
        red[0] += part_bb[k];
        red[1] += part_rr[k];
    }
    block_sum_n<2>(red, s_w);
    if (threadIdx.x == 0) {
        const double bTb = red[0], rho = red[1];
        const bool zero_rhs = bTb < 1.0e-10;
        const bool zero_rho = !zero_rhs && rho < 1.0e-10;
        sc->bTb = bTb;
        sc->criteria = bTb * crit_factor;          // 1e-8 in the reference
        sc->rho_s[0] = 1.0;
        sc->rho_s[1] = rho;              // rho of iteration 1
        sc->alpha = 1.0;
        sc->omega = 1.0;
        sc->rTr = 0.0;
        sc->iter = 0;
        sc->done = (zero_rhs || zero_rho) ? 1 : 0;
        sc->kind = zero_rhs ? CAPAC_SOLVE_ZERO_RHS : (zero_rho ? CAPAC_SOLVE_ZERO_RHO : CAPAC_SOLVE_CONVERGED);
    }
}

// K2 of iteration n.  For n >= 2 it first completes iteration n-1: rTr from
// the (r, r) partials of K4, the stop test and the 1000-iteration rule
// (bicg_stab.c:539-571); then alpha = rho_n / (r0, v), s = r - alpha v, z = K s
// (bicg_stab.c:416-460)
__global__ void __launch_bounds__(kVecThreads)
k_vec_s(int n_mat, int n, int n_groups, const double* __restrict__ part1, int n_vb,
        const double* __restrict__ part4, const double* __restrict__ r, const double* __restrict__ v,
        const double* __restrict__ K, const double* __restrict__ r0, double* __restrict__ s,
        double* __restrict__ z, double* __restrict__ part2, SolveScalars* sc) {
    main_stream_priority();
    __shared__ double s_w[16 * 2];
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = sc->done;
    // this thread's row is loaded before the reductions (it does not depend
    // on them), so the two memory round trips overlap
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double rk = 0.0, vk = 0.0, Kk = 0.0, r0k = 0.0;
    if (k < n_mat) { rk = r[k]; vk = v[k]; Kk = K[k]; r0k = r0[k]; }
    __syncthreads();
    if (s_skip) return;
    double red[2] = {0.0, 0.0};          // (r0, v), (r, r) of iteration n-1
    for (int k = threadIdx.x; k < n_groups; k += blockDim.x) red[0] += part1[k];
    if (n >= 2)
        for (int k = threadIdx.x; k < n_vb; k += blockDim.x) red[1] += part4[k];
    block_sum_n<2>(red, s_w);
    if (n >= 2) {
        const double rTr = red[1];
        const int iter = n - 1;
        const double crit = sc->criteria;
        int done = 0, kind = CAPAC_SOLVE_CONVERGED;
        if (rTr < crit) {
            done = 1;
        } else if (iter > 1000) {
            done = 1;
            kind = rTr < crit * 100.0 ? CAPAC_SOLVE_LOOSENED : CAPAC_SOLVE_FAILED;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->iter = iter;
            sc->rTr = rTr;
            if (done) { sc->kind = kind; sc->done = 1; }
        }
        if (done) return;
    }
    const double r0v = red[0];
    const double alpha = sc->rho_s[n & 1] / r0v;
    double acc = 0.0;
    if (k < n_mat) {
        double sk = rk - alpha * vk;
        s[k] = sk;
        z[k] = sk * Kk;
        acc = r0k * sk;
    }
    __syncthreads();                      // s_w is reused
    double r0s = block_sum(acc, s_w);
    if (threadIdx.x == 0) {
        part2[blockIdx.x] = r0s;
        if (blockIdx.x == 0) sc->alpha = alpha;
    }
}

// K4 of iteration n: omega, x += alpha y + omega z, r = s - omega t
// (bicg_stab.c:483-535), then rho_{n+1}, beta and the next direction
// p = r + beta (p - omega v), y = K p (bicg_stab.c:365-404)
__global__ void __launch_bounds__(kVecThreads)
k_vec_xp(int n_mat, int n, int n_groups, const double* __restrict__ part3, int n_vb,
         const double* __restrict__ part2, double* __restrict__ x, double* __restrict__ y,
         const double* __restrict__ z, const double* __restrict__ s, const double* __restrict__ t,
         double* __restrict__ r, double* __restrict__ p, const double* __restrict__ v,
         const double* __restrict__ K, double* __restrict__ part4, SolveScalars* sc) {
    main_stream_priority();
    __shared__ double s_w[16 * 4];
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = sc->done;
    // this thread's row is loaded before the reductions (see k_vec_s)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double xk = 0.0, yk = 0.0, zk = 0.0, sk = 0.0, tk = 0.0, pk0 = 0.0, vk = 0.0, Kk = 0.0;
    if (k < n_mat) {
        xk = x[k]; yk = y[k]; zk = z[k]; sk = s[k]; tk = t[k]; pk0 = p[k]; vk = v[k]; Kk = K[k];
    }
    __syncthreads();
    if (s_skip) return;
    double red[4] = {0.0, 0.0, 0.0, 0.0};   // (Kt, z), (Kt, Kt), (r0, t), (r0, s)
    for (int k = threadIdx.x; k < n_groups; k += blockDim.x) {
        red[0] += part3[3 * k];
        red[1] += part3[3 * k + 1];
        red[2] += part3[3 * k + 2];
    }
    for (int k = threadIdx.x; k < n_vb; k += blockDim.x) red[3] += part2[k];
    block_sum_n<4>(red, s_w);
    const double ktz = red[0], ktkt = red[1], r0t = red[2], r0s = red[3];
    const double omega = ktz / ktkt;
    const double alpha = sc->alpha;
    const double rho = sc->rho_s[n & 1];
    const double rho_next = r0s - omega * r0t;
    const double beta = (rho_next / rho) * (alpha / omega);
    double acc = 0.0;
    if (k < n_mat) {
        x[k] = xk + (alpha * yk + omega * zk);
        double rk = sk - omega * tk;
        r[k] = rk;
        acc = rk * rk;
        double pk = rk + beta * (pk0 - omega * vk);
        p[k] = pk;
        y[k] = pk * Kk;
    }
    __syncthreads();                      // s_w is reused
    double rr = block_sum(acc, s_w);
    if (threadIdx.x == 0) {
        part4[blockIdx.x] = rr;
        if (blockIdx.x == 0) {
            sc->rho_s[(n + 1) & 1] = rho_next;
            sc->omega = omega;
        }
    }
}

// ------------------------------------------------------------------ GMRES
// The FP32-refined path's inner solver (the reference's BiCGSTAB stays the
// FP64 path's): GMRES(kGmM) right-preconditioned with the same Jacobi K, so
// it minimises |b - A (x0 + K V y)| over the Krylov space of A K.  The relay
// is symmetric but indefinite, and its signed Jacobi preconditioner is what
// makes it converge fast; GMRES keeps that preconditioner and needs one
// product per column where BiCGSTAB needs two per iteration (Au561 and Au2057
// decks from the quadratic warm start, dense oracle relay: 14 against 23 and
// 19 against 29 products to rTr < 1e-8 bTb).  Column j is one unit:
//   SpMV   w = A u, u = K q_j (q_j unnormalised)
//   dots   partials of (w, q_i), i <= j (a second grid dimension over groups
//          of basis vectors); one workgroup also forms |q_j| from the
//          partials that formed q_j, rotates column j - 1 and takes the stop
//          test (so one product past the stop is spent, as BiCGSTAB's
//          deferred test does)
//   update w /= |q_j| and v_j = q_j / |q_j| in place; h_i from the partials;
//          q_{j+1} = w - sum h_i v_i (classical Gram-Schmidt), u = K q_{j+1},
//          partials of |q_{j+1}|^2
// x is updated once, by k_gm_finish, from y = R^-1 g.  Every block reduces
// the partials itself in the same fixed order, as the vector kernels do.
// The stop test is the reference's rTr < crit bTb on GMRES's residual
// estimate |g_{j+1}|^2; the 1000-iteration rule counts columns.

// start: r0 = b - Ax (Ax formed by the product before), q_0 = r0, K, u = K r0,
// partials of (b, b) and (r0, r0)
__global__ void __launch_bounds__(kVecThreads)
k_gm_init(int n_mat, int n4, const double* __restrict__ b, const double* __restrict__ Ax,
          const double* __restrict__ diag, double* __restrict__ q0, double* __restrict__ u,
          double* __restrict__ K, double* __restrict__ part_bb, double* __restrict__ part_qq) {
    main_stream_priority();
    __shared__ double s_w[16 * 2];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (k < n_mat) {
        const double bk = b[k];
        const double rk = bk - Ax[k];
        const double Kk = k < n4 ? 1.0 / diag[k] : 1.0;
        q0[k] = rk;
        K[k] = Kk;
        u[k] = Kk * rk;
        acc[0] = bk * bk;
        acc[1] = rk * rk;
    }
    block_sum_n<2>(acc, s_w);
    if (threadIdx.x == 0) {
        part_bb[blockIdx.x] = acc[0];
        part_qq[blockIdx.x] = acc[1];
    }
}

// one workgroup: bTb, criteria, |r0|; the early exits of a first start (x
// unchanged): bTb < 1e-10, and (r0, r0) < 1e-10 (BiCGSTAB's first rho); a
// restart whose residual already meets the test ends converged
__global__ void __launch_bounds__(256)
k_gm_start(int n_vb, const double* __restrict__ part_bb, const double* __restrict__ part_qq,
           SolveScalars* sc, GmState* gm, double crit_factor, int first) {
    main_stream_priority();
    __shared__ double s_w[16 * 2];
    double red[2] = {0.0, 0.0};
    for (int k = threadIdx.x; k < n_vb; k += blockDim.x) {
        red[0] += part_bb[k];
        red[1] += part_qq[k];
    }
    block_sum_n<2>(red, s_w);
    if (threadIdx.x != 0) return;
    const double bTb = red[0], rr = red[1];
    sc->bTb = bTb;
    sc->criteria = bTb * crit_factor;
    sc->rTr = rr;
    if (first) sc->iter = 0;
    gm->nrm[0] = sqrt(rr);
    gm->g[0] = sqrt(rr);
    gm->ny = 0;
    if (first && bTb < 1.0e-10) {
        sc->done = 1; sc->kind = CAPAC_SOLVE_ZERO_RHS;
    } else if (first && rr < 1.0e-10) {
        sc->done = 1; sc->kind = CAPAC_SOLVE_ZERO_RHO;
    } else if (!first && rr < sc->criteria) {
        sc->done = 1; sc->kind = CAPAC_SOLVE_CONVERGED;
    } else {
        sc->done = 0; sc->kind = CAPAC_SOLVE_CONVERGED;
    }
}

// kind of a solve that stopped to restart from its updated x (host-side only)
constexpr int kSolveRestart = 5;

// one workgroup (every thread calls it): rotate column jj (its h_{jj+1,jj} =
// hnext), the residual estimate and the stop test; when the solve stops or
// the basis is full, the update's coefficients y = R^-1 g.  The column, the
// rotations and R are staged in LDS, so the serial work of thread 0 touches
// no global memory.
struct GmLds {
    double col[kGmM + 1], cs[kGmM], sn[kGmM], g[kGmM + 1], y[kGmM];
    double R[kGmM][kGmM];       // column l of R in R[l][0..l]
    int stop;
};
__device__ void gm_column(GmState* gm, SolveScalars* sc, int jj, double hnext, GmLds& L) {
    const int t = threadIdx.x;
    if (t <= jj) L.col[t] = gm->H[jj][t];
    if (t < jj) { L.cs[t] = gm->cs[t]; L.sn[t] = gm->sn[t]; }
    if (t == 0) L.g[jj] = gm->g[jj];
    __syncthreads();
    if (t == 0) {
        double* col = L.col;
        for (int i = 0; i < jj; ++i) {
            const double u = L.cs[i] * col[i] + L.sn[i] * col[i + 1];
            col[i + 1] = -L.sn[i] * col[i] + L.cs[i] * col[i + 1];
            col[i] = u;
        }
        const double a = col[jj];
        const double rad = sqrt(a * a + hnext * hnext);
        const double c = rad > 0.0 ? a / rad : 1.0, sn = rad > 0.0 ? hnext / rad : 0.0;
        col[jj] = rad;
        const double gj = L.g[jj];
        const double g1 = -sn * gj;
        gm->cs[jj] = c;
        gm->sn[jj] = sn;
        gm->g[jj + 1] = g1;
        gm->g[jj] = c * gj;
        const double res2 = g1 * g1;
        sc->rTr = res2;
        sc->iter += 1;
        const bool converged = res2 < sc->criteria;
        const bool over = sc->iter > 1000;
        L.stop = converged || over || jj + 1 == kGmM;
        if (L.stop) {
            sc->done = 1;
            if (converged) sc->kind = CAPAC_SOLVE_CONVERGED;
            else if (over) sc->kind = res2 < 100.0 * sc->criteria ? CAPAC_SOLVE_LOOSENED : CAPAC_SOLVE_FAILED;
            else sc->kind = kSolveRestart;
        }
    }
    __syncthreads();
    if (t <= jj) gm->H[jj][t] = L.col[t];
    if (!L.stop) return;
    // y = R^-1 g over columns 0 .. jj, R and g staged first
    for (int e = t; e < (jj + 1) * (jj + 1); e += blockDim.x) {
        const int l = e / (jj + 1), i = e % (jj + 1);
        L.R[l][i] = l == jj ? L.col[i] : (i <= l ? gm->H[l][i] : 0.0);
    }
    if (t <= jj) L.g[t] = gm->g[t];     // thread 0's writes above are visible after the barrier
    __syncthreads();
    if (t == 0) {
        for (int i = jj; i >= 0; --i) {
            double u = L.g[i];
            for (int l = i + 1; l <= jj; ++l) u -= L.R[l][i] * L.y[l];
            L.y[i] = L.R[i][i] != 0.0 ? u / L.R[i][i] : 0.0;
        }
        gm->ny = jj + 1;
    }
    __syncthreads();
    if (t <= jj) gm->y[t] = L.y[t];
}

// column j's dot products (see the block comment above), grid (row blocks,
// groups of kGmDots basis vectors): partials of (w, q_i) for the product as it
// came, with q_0 .. q_{j-1} stored normalised (the update launch normalises
// q_j when it has |q_j|); the update launch applies the 1 / |q_j| scales, so
// no workgroup here waits for |q_j| except workgroup (0, 0), which forms it,
// keeps it and rotates column j - 1.  j == kGmM only rotates the last column.
// With qcopy, the workgroups that load q_j also copy it there: the block
// update launch reads q_j from the copy, since its workgroups normalise Q[j]
// in place while others of the same block still read it.
constexpr int kGmDots = 8;
__global__ void __launch_bounds__(kVecThreads)
k_gm_dots(int n_mat, int j, const double* __restrict__ w, const double* __restrict__ Q,
          const double* __restrict__ part_qq, int n_vb, GmState* gm, SolveScalars* sc,
          double* __restrict__ part_h, double* __restrict__ qcopy) {
    main_stream_priority();
    __shared__ double s_w[16 * kGmDots];
    __shared__ GmLds s_gm;
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = sc->done;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i0 = blockIdx.y * kGmDots;
    const bool row = k < n_mat && j < kGmM;
    const bool lead = blockIdx.x == 0 && blockIdx.y == 0;
    // every load of the launch before the first wait
    const double wraw = row ? w[k] : 0.0;
    double qv[kGmDots];
#pragma unroll
    for (int t = 0; t < kGmDots; ++t) qv[t] = row && i0 + t <= j ? Q[(size_t)(i0 + t) * n_mat + k] : 0.0;
    if (qcopy && row && i0 <= j && j < i0 + kGmDots) {
#pragma unroll
        for (int t = 0; t < kGmDots; ++t)
            if (i0 + t == j) qcopy[k] = qv[t];
    }
    double pq = 0.0;
    if (lead)
        for (int b = threadIdx.x; b < n_vb; b += blockDim.x) pq += part_qq[b];
    __syncthreads();
    if (s_skip) return;
    if (lead) {
        const double nrm_j = sqrt(block_sum(pq, s_w));
        if (threadIdx.x == 0) gm->nrm[j] = nrm_j;
        if (j >= 1) gm_column(gm, sc, j - 1, nrm_j, s_gm);
        __syncthreads();                  // s_w is reused
    }
    if (j >= kGmM) return;
    double d[kGmDots];
#pragma unroll
    for (int t = 0; t < kGmDots; ++t) d[t] = wraw * qv[t];
    block_sum_n<kGmDots>(d, s_w);
    if (threadIdx.x == 0)
#pragma unroll
        for (int t = 0; t < kGmDots; ++t)
            if (i0 + t <= j) part_h[(size_t)blockIdx.x * (kGmM + 1) + i0 + t] = d[t];
}

// Column j of the block-preconditioned GMRES on the TILES layout: the
// product's sums launch and the dots launch in one.  Workgroup b < Nb forms
// row block b of w = A u from the tile partials (as k_spmv_tsum) and then
// the partial dots (w, q_i), i <= j, over its 256 rows into part_h[b], and
// copies q_j's rows into qcopy (the update launch reads q_j there);
// workgroup Nb + k takes the Lagrange row of nanoparticle k (formed here
// when the tile launch left it to the sums: lag_here) and its dots into
// part_h[Nb + k]; the last workgroup does what the dots launch's workgroup
// (0, 0) does besides its dots (|q_j| from part_qq, the rotation of column
// j - 1, the stop test).  The update launch then reduces Nb + n_np rows of
// part_h.
__global__ void __launch_bounds__(kSumThreads)
k_gm_tsum_dots(int Nb, int Nm, int n_np, int n_mat, int j, const int* __restrict__ np_of,
               const int* __restrict__ np_first, const int* __restrict__ np_last, const double* __restrict__ P,
               BoxConst bc, const double* __restrict__ x, int lag_here, double* __restrict__ y,
               const double* __restrict__ Q, double* __restrict__ qcopy, const double* __restrict__ part_qq,
               int n_qq, GmState* gm, SolveScalars* sc, double* __restrict__ part_h) {
    main_stream_priority();
    __shared__ double s_o[kSumThreads / 64][4][kTileEdge];       // [group][component][row]
    __shared__ double s_y[4][kTileEdge];                         // the block's outputs
    __shared__ double s_w[(kSumThreads / 64) * kGmDots];
    __shared__ GmLds s_gm;
    if (sc->done) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int b = blockIdx.x;
    if (b == Nb + n_np) {
        double pq = 0.0;
        for (int k = t; k < n_qq; k += blockDim.x) pq += part_qq[k];
        const double nrm_j = sqrt(block_sum(pq, s_w));
        if (t == 0) gm->nrm[j] = nrm_j;
        if (j >= 1) gm_column(gm, sc, j - 1, nrm_j, s_gm);
        return;
    }
    int row = -1;
    double wr = 0.0;
    if (b < Nb) {
        // the row sums over the source blocks, as tile_block_out
        const int G = kSumThreads >> 6, g = wave;
        const double* src = P + (size_t)b * Nb * (4 * kTileEdge) + lane;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int q0 = g; q0 < Nb; q0 += G * kSumBatch) {
            double v[kSumBatch][4];
#pragma unroll
            for (int u = 0; u < kSumBatch; ++u) {
                const int q = q0 + G * u;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    v[u][c] = q < Nb ? src[(size_t)q * (4 * kTileEdge) + c * kTileEdge] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kSumBatch; ++u)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] += v[u][c];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) s_o[g][c][lane] = acc[c];
        __syncthreads();
        const int ia = b * kTileEdge + lane;
        if (t < kTileEdge) {
            double o[4] = {0.0, 0.0, 0.0, 0.0};
            if (ia < Nm) {
                for (int h = 0; h < G; ++h)
#pragma unroll
                    for (int c = 0; c < 4; ++c) o[c] += s_o[h][c][t];
                RelayG32 m{NULL, 0, NULL, NULL, bc};
                spmv_group_out(m, ia, Nm, np_of, x, y, 0, NULL, NULL, NULL, o);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) s_y[c][t] = o[c];
        }
        __syncthreads();
        if (ia < Nm) {
            const int c = wave;
            row = c < 3 ? 3 * ia + c : 3 * Nm + ia;
            wr = s_y[c][lane];
        }
    } else {
        // Lagrange row of nanoparticle k
        const int k = b - Nb;
        if (lag_here) {
            double v = 0.0;
            for (int mi = np_first[k] + t; mi <= np_last[k]; mi += kSumThreads) v += x[3 * Nm + mi];
            v = block_sum(v, s_w);
            __syncthreads();                 // s_w is reused below
            if (t == 0) y[4 * Nm + k] = v;
            wr = v;
        } else {
            wr = y[4 * Nm + k];
        }
        if (t == 0) row = 4 * Nm + k;
    }
    // the partial dots, kGmDots basis vectors per round
    for (int i0 = 0; i0 <= j; i0 += kGmDots) {
        double d[kGmDots];
#pragma unroll
        for (int u = 0; u < kGmDots; ++u) {
            const int i = i0 + u;
            const double qv = row >= 0 && i <= j ? Q[(size_t)i * n_mat + row] : 0.0;
            if (i == j && row >= 0) qcopy[row] = qv;
            d[u] = wr * qv;
        }
#pragma unroll
        for (int u = 0; u < kGmDots; ++u) d[u] = wave_sum(d[u]);
        if (lane == 0)
#pragma unroll
            for (int u = 0; u < kGmDots; ++u) s_w[wave * kGmDots + u] = d[u];
        __syncthreads();
        if (t < kGmDots && i0 + t <= j) {
            double v = 0.0;
            for (int w = 0; w < kSumThreads / 64; ++w) v += s_w[w * kGmDots + t];
            part_h[(size_t)b * (kGmM + 1) + i0 + t] = v;
        }
        __syncthreads();                     // s_w is reused by the next round
    }
}

// column j's Gram-Schmidt step: w = w / |q_j| and v_j = q_j / |q_j| in place,
// h_i from the partials (scaled by 1 / |q_j|, and 1 / |q_j|^2 for i = j),
// q_{j+1} = w - sum h_i v_i, u = K q_{j+1}, partials of |q_{j+1}|^2; block 0
// keeps the column's h.  The basis entries of the row are loaded before the
// reduction of h, so the launch waits for memory once.
__global__ void __launch_bounds__(kVecThreads)
k_gm_update(int n_mat, int j, double* __restrict__ w, double* __restrict__ Q,
            const double* __restrict__ K, const double* __restrict__ part_h, int n_vb, GmState* gm,
            const SolveScalars* __restrict__ sc, double* __restrict__ u, double* __restrict__ part_qq) {
    main_stream_priority();
    __shared__ double s_w[16];
    __shared__ double s_h[kGmM + 1];
    __shared__ double s_g[kVecThreads / (kGmM + 1)][kGmM + 1];
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = sc->done;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double wk = 0.0, Kk = 0.0, qj = 0.0;
    double qv[kGmM];
    if (k < n_mat) { wk = w[k]; Kk = K[k]; qj = Q[(size_t)j * n_mat + k]; }
#pragma unroll
    for (int i = 0; i < kGmM; ++i) qv[i] = k < n_mat && i < j ? Q[(size_t)i * n_mat + k] : 0.0;
    // h_i = sum over the blocks' partials: thread (g, i) adds blocks g, g + G,
    // ... (G = 6 groups, loads in flight together), then the groups in order
    constexpr int G = kVecThreads / (kGmM + 1);
    const int ih = threadIdx.x % (kGmM + 1), g = threadIdx.x / (kGmM + 1);
    double t = 0.0;
    if (g < G && ih <= j)
        for (int b0 = g; b0 < n_vb; b0 += 8 * G) {
            double pv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                pv[q] = b0 + q * G < n_vb ? part_h[(size_t)(b0 + q * G) * (kGmM + 1) + ih] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) t += pv[q];
        }
    __syncthreads();
    if (s_skip) return;
    if (g < G) s_g[g][ih] = t;
    __syncthreads();
    const double inv = 1.0 / gm->nrm[j];
    if (threadIdx.x <= j) {
        double h = 0.0;
        for (int gg = 0; gg < G; ++gg) h += s_g[gg][threadIdx.x];
        s_h[threadIdx.x] = (int)threadIdx.x == j ? h * inv * inv : h * inv;
    }
    __syncthreads();
    double acc = 0.0;
    if (k < n_mat) {
        const double vj = qj * inv;
        Q[(size_t)j * n_mat + k] = vj;
        const double wn = wk * inv;
        w[k] = wn;
        double q = wn;
#pragma unroll
        for (int i = 0; i < kGmM; ++i)
            if (i < j) q -= s_h[i] * qv[i];
        q -= s_h[j] * vj;
        Q[(size_t)(j + 1) * n_mat + k] = q;
        u[k] = Kk * q;
        acc = q * q;
    }
    const double qq = block_sum(acc, s_w);
    if (threadIdx.x == 0) part_qq[blockIdx.x] = qq;
    if (blockIdx.x == 0 && threadIdx.x <= j) gm->H[j][threadIdx.x] = s_h[threadIdx.x];
}

// Column j's Gram-Schmidt step with the block preconditioner, aligned to its
// blocks (precond.h): workgroup (b, s) forms q_{j+1} on all 256 unknowns of
// block b as k_gm_update forms it on its rows, writes quarter s of them (v_j,
// w and q_{j+1}) with its partial of |q_{j+1}|^2, and sets u = B_b^-1 q_{j+1}
// on the quarter from the block's inverse (k_bj_apply's sums); quarter 3
// also keeps Z_b^T q_{j+1} in cs for the coarse correction, which the
// product adds.  The last workgroup takes the Lagrange rows (preconditioner
// 1).  Partials go to part_qq[g] for workgroup g (zero from idle ones).
// Every workgroup of a block reads w and q_j on all of the block's rows, and
// the owners of the other quarters write Q[j] in the same launch, so q_j
// comes from the dots launch's copy qj_in, and w (not read after this
// launch) is left unnormalised.
__global__ void __launch_bounds__(kVecThreads)
k_gm_update_bj(int n_mat, int Nm, int Nb, int j, const double* __restrict__ w, const double* __restrict__ qj_in,
               double* __restrict__ Q,
               const double* __restrict__ part_h, int n_vb, GmState* gm, const SolveScalars* __restrict__ sc,
               double* __restrict__ u, double* __restrict__ part_qq, const float* __restrict__ binv,
               const float4* __restrict__ cw, double* __restrict__ cs) {
    main_stream_priority();
    __shared__ double s_w[16];
    __shared__ double s_h[kGmM + 1];
    __shared__ double s_g[kVecThreads / (kGmM + 1)][kGmM + 1];
    __shared__ double s_q[kBjUnk];
    __shared__ double s_red[4][kTileEdge];
    __shared__ int s_skip;
    const int g = blockIdx.x, t = threadIdx.x;
    const bool lag = g == bj_groups(Nb);
    int b = 0, s = 0;
    if (!lag) bj_group(g, b, s);
    if (!lag && b >= Nb) {
        if (t == 0) part_qq[g] = 0.0;
        return;
    }
    if (t == 0) s_skip = sc->done;
    // this thread's unknown: rho = t of block b, or Lagrange row 4 Nm + t
    const int row = lag ? (4 * Nm + t < n_mat ? 4 * Nm + t : -1) : bj_row(b, t, Nm);
    const bool mine = row >= 0 && (lag || (t >> 6) == s);      // the rows this workgroup writes
    double wk = 0.0, qj = 0.0;
    double qv[kGmM];
    if (row >= 0) { wk = w[row]; qj = qj_in[row]; }
#pragma unroll
    for (int i = 0; i < kGmM; ++i) qv[i] = row >= 0 && i < j ? Q[(size_t)i * n_mat + row] : 0.0;
    constexpr int G = kVecThreads / (kGmM + 1);
    const int ih = t % (kGmM + 1), gg0 = t / (kGmM + 1);
    double hs = 0.0;
    if (gg0 < G && ih <= j)
        for (int b0 = gg0; b0 < n_vb; b0 += 8 * G) {
            double pv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                pv[q] = b0 + q * G < n_vb ? part_h[(size_t)(b0 + q * G) * (kGmM + 1) + ih] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) hs += pv[q];
        }
    __syncthreads();
    if (s_skip) return;
    if (gg0 < G) s_g[gg0][ih] = hs;
    __syncthreads();
    const double inv = 1.0 / gm->nrm[j];
    if (t <= j) {
        double h = 0.0;
        for (int q = 0; q < G; ++q) h += s_g[q][t];
        s_h[t] = t == j ? h * inv * inv : h * inv;
    }
    __syncthreads();
    double q = 0.0;
    if (row >= 0) {
        const double vj = qj * inv;
        const double wn = wk * inv;
        q = wn;
#pragma unroll
        for (int i = 0; i < kGmM; ++i)
            if (i < j) q -= s_h[i] * qv[i];
        q -= s_h[j] * vj;
        if (mine) {
            Q[(size_t)j * n_mat + row] = vj;
            Q[(size_t)(j + 1) * n_mat + row] = q;
        }
    }
    if (g == 0 && t <= j) gm->H[j][t] = s_h[t];
    if (lag) {
        if (row >= 0) u[row] = q;
        const double qq = block_sum(row >= 0 ? q * q : 0.0, s_w);
        if (t == 0) part_qq[g] = qq;
        return;
    }
    if ((t >> 6) == s) {
        const double qq = wave_sum(mine ? q * q : 0.0);
        if ((t & 63) == 0) part_qq[g] = qq;
    }
    s_q[t] = q;
    if (cs && s == 3 && t >= 3 * kTileEdge) {
        // the block's coarse components Z_b^T q (its charges are unknowns 192..255)
        const float4 z = cw[b * kTileEdge + (t & 63)];
        const double c0 = wave_sum(z.x * q), c1 = wave_sum(z.y * q), c2 = wave_sum(z.z * q), c3 = wave_sum(z.w * q);
        if ((t & 63) == 0) {
            cs[4 * b] = c0; cs[4 * b + 1] = c1; cs[4 * b + 2] = c2; cs[4 * b + 3] = c3;
        }
    }
    __syncthreads();
    // u = B_b^-1 q on quarter s: thread (ch, l) sums columns 64 ch .. 64 ch + 63 of row 64 s + l
    const int ro = s * kTileEdge + (t & 63), ch = t >> 6;
    const float4* mp = reinterpret_cast<const float4*>(binv + (size_t)b * kBjUnk * kBjUnk) + ro;
    float4 mv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) mv[i] = mp[(size_t)(16 * ch + i) * kBjUnk];
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
    const int orow = bj_row(b, ro, Nm);
    if (orow >= 0) u[orow] = ((s_red[0][t] + s_red[1][t]) + s_red[2][t]) + s_red[3][t];
}

// x += K sum_i y_i v_i once the solve (or a restart) is due; no-op otherwise.
// With t_out, sum_i y_i v_i goes there instead (zeros when y is empty), for
// the block preconditioner's application.
__global__ void __launch_bounds__(kVecThreads)
k_gm_finish(int n_mat, const double* __restrict__ Q, const double* __restrict__ K, const GmState* gm,
            const SolveScalars* __restrict__ sc, double* __restrict__ x, double* __restrict__ t_out) {
    main_stream_priority();
    if (!sc->done) return;
    const int ny = gm->ny;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (t_out && ny == 0 && k < n_mat) t_out[k] = 0.0;
    if (ny == 0 || k >= n_mat) return;
    double t = 0.0;
    for (int i0 = 0; i0 < ny; i0 += 8) {
        double vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) vv[u] = i0 + u < ny ? Q[(size_t)(i0 + u) * n_mat + k] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u < ny) t += gm->y[i0 + u] * vv[u];
    }
    if (t_out) t_out[k] = t;
    else x[k] += K[k] * t;
}

// (rr, rr) of the FP64 residual from its per-group partials (with a guard:
// only after a solve that finished converged or loosened)
__global__ void __launch_bounds__(256)
k_res_finish(const double* __restrict__ part, int n, double* __restrict__ res,
             const SolveScalars* __restrict__ guard) {
    main_stream_priority();
    __shared__ double s_w[16];
    if (guard && !(guard->done && (guard->kind == CAPAC_SOLVE_CONVERGED || guard->kind == CAPAC_SOLVE_LOOSENED)))
        return;
    double v = block_reduce(part, n, 1, 0, s_w);
    if (threadIdx.x == 0) res[0] = v;
}

__global__ void __launch_bounds__(256) k_axpy(int n, double a, const double* __restrict__ d, double* __restrict__ x) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] += a * d[k];
}

__global__ void __launch_bounds__(256) k_scale(int n, double a, double* __restrict__ x) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] *= a;
}

__global__ void __launch_bounds__(256) k_zero(int n, double* __restrict__ x) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = 0.0;
}

// The start of a force call in one launch: red[0..n_red) cleared, the warm
// start extrapolated from the last solutions when an MD step asked for it
// (h0 newest; order 1: 2 h0 - h1, order 2: 3 h0 - 3 h1 + h2), and x kept in
// x_backup for a call that has to be repeated
__global__ void __launch_bounds__(256)
k_force_prologue(int n, int order, const double* __restrict__ h0, const double* __restrict__ h1,
                 const double* __restrict__ h2, double* __restrict__ x, double* __restrict__ x_backup,
                 double* __restrict__ red, int n_red) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_red) red[k] = 0.0;
    if (k >= n) return;
    double xk;
    if (order == 1) { xk = 2.0 * h0[k] - h1[k]; x[k] = xk; }
    else if (order == 2) { xk = 3.0 * (h0[k] - h1[k]) + h2[k]; x[k] = xk; }
    else xk = x[k];
    x_backup[k] = xk;
}

}  // namespace

int solver_create(capac_ctx* c) {
    Solver* S = (Solver*)calloc(1, sizeof(Solver));
    if (!S) { capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
    c->solver = S;
    S->n_mat = c->n_mat;
    S->Nm = c->bc.Nm;
    S->n_np = c->bc.n_np;
    S->n_groups = S->Nm + S->n_np;
    S->n_vb = (S->n_mat + kVecThreads - 1) / kVecThreads;
    size_t n = (size_t)(c->n_mat > 0 ? c->n_mat : 1);
    double** vecs[12] = {&S->r0, &S->r, &S->p, &S->v, &S->s, &S->t, &S->y, &S->z, &S->K, &S->Ax,
                         &S->d, &S->rr};
    for (int k = 0; k < 12; ++k) {
        CHECK_HIP(hipMalloc(vecs[k], n * sizeof(double)));
        CHECK_HIP(hipMemset(*vecs[k], 0, n * sizeof(double)));
    }
    const size_t ng = (size_t)S->n_groups + 1, nvb = (size_t)S->n_vb + 1;
    CHECK_HIP(hipMalloc(&S->part1, ng * sizeof(double)));
    CHECK_HIP(hipMalloc(&S->part3, 3 * ng * sizeof(double)));
    CHECK_HIP(hipMalloc(&S->part_res, ng * sizeof(double)));
    CHECK_HIP(hipMalloc(&S->part2, nvb * sizeof(double)));
    // part4 also takes the block-aligned update's partials (one per workgroup)
    const size_t n4p = std::max(nvb, (size_t)bj_groups(tile_blocks(S->Nm)) + 2);
    CHECK_HIP(hipMalloc(&S->part4, n4p * sizeof(double)));
    if (c->precision == CAPAC_PREC_FP32_REFINED) {
        CHECK_HIP(hipMalloc(&S->Q, (size_t)(kGmM + 1) * n * sizeof(double)));
        CHECK_HIP(hipMemset(S->Q, 0, (size_t)(kGmM + 1) * n * sizeof(double)));
        S->ph_rows = (int)std::max(nvb, (size_t)tile_blocks(S->Nm) + (size_t)S->n_np);
        CHECK_HIP(hipMalloc(&S->part_h, (size_t)S->ph_rows * (kGmM + 1) * sizeof(double)));
        CHECK_HIP(hipMalloc(&S->gm, sizeof(GmState)));
        CHECK_HIP(hipMemset(S->gm, 0, sizeof(GmState)));
    }
    S->sc = (SolveScalars*)(c->red + kRedSc);
    S->sc_host = (SolveScalars*)(c->red_host + kRedSc);
    S->res = c->red + kRedRes;
    S->res_host = c->red_host + kRedRes;
    S->in_red = true;
    return 0;
}

void solver_destroy(capac_ctx* c) {
    Solver* S = c->solver;
    if (!S) return;
    void* vecs[] = {S->r0, S->r, S->p, S->v, S->s, S->t, S->y, S->z, S->K, S->Ax, S->d, S->rr,
                    S->part1, S->part3, S->part_res, S->part2, S->part4, S->Q, S->part_h, S->gm};
    for (size_t k = 0; k < sizeof vecs / sizeof vecs[0]; ++k)
        if (vecs[k]) (void)hipFree(vecs[k]);
    free(S);
    c->solver = NULL;
}

template <class M, int R>
static void spmv_t(capac_ctx* c, const Solver* S, const M& m, const double* x, double* y, int mode,
                   double* part, const SolveScalars* sc) {
    const int blocks = (S->Nm + R - 1) / R + S->n_np;
    k_spmv<M, R><<<blocks, 256, 0, c->stream>>>(m, S->Nm, S->n_np, c->np_of_metal, c->np_first,
                                                c->np_last, x, y, mode, S->r0, S->K, part, sc);
}

// Rows per workgroup: the FP64 product is VALU work (every block formed
// from the positions) and takes one row-group per workgroup.  The FP32
// product streams 16 bytes per pair; several row-groups per workgroup share
// each column's vector entries and position, which the lanes otherwise load
// once per pair, while small decks keep one row-group per workgroup so that
// there are enough workgroups to fill the chip.
#ifndef CAPAC_G32_ROWS_LARGE
#define CAPAC_G32_ROWS_LARGE 2
#endif
#ifndef CAPAC_G32_ROWS_MID
#define CAPAC_G32_ROWS_MID 2
#endif
#ifndef CAPAC_G32_LARGE_MIN
#define CAPAC_G32_LARGE_MIN 1536
#endif
#ifndef CAPAC_G32_MID_MIN
#define CAPAC_G32_MID_MIN 768
#endif
// coarse (tiles only): x is B^-1 q from the block-preconditioned update and
// the product adds the coarse correction first (CoarseArgs), its corrected
// input kept in S->t
static void spmv(capac_ctx* c, const Solver* S, const double* x, double* y, int mode, double* part,
                 const SolveScalars* sc, bool coarse = false) {
    if (c->precision == CAPAC_PREC_FP32_REFINED && c->relay_layout == CAPAC_LAYOUT_TILES) {
        const int nb = c->relay_nb, nt = (int)tile_count(nb);
        CoarseArgs cg{NULL, NULL, NULL, NULL, 0};
        if (coarse && c->cg_n) cg = CoarseArgs{c->cg_A, c->cg_s, reinterpret_cast<const float4*>(c->cg_w), S->t,
                                               c->cg_n};
        // the Lagrange rows: in the tile launch, or with a coarse correction in the sums launch
        const int lag_t = cg.nc ? 0 : S->n_np, lag_s = cg.nc ? S->n_np : 0;
        k_spmv_tiles<<<nt + lag_t, kTileThreads, 0, c->stream>>>(
            reinterpret_cast<const float4*>(c->relay32), nb, nt, S->Nm, S->n_np, c->np_of_metal, c->np_first,
            c->np_last, c->r, c->metal_u, c->bc, x, y, c->relay_part, c->relay_cnt, mode, S->r0, S->K, part,
            sc, cg);
        if (!CAPAC_TILE_FUSED)
            k_spmv_tsum<<<nb + lag_s, kSumThreads, 0, c->stream>>>(nb, S->Nm, c->np_of_metal, c->relay_part, c->bc,
                                                                   cg.nc ? cg.x2 : x, y, mode, S->r0, S->K, part,
                                                                   sc, c->np_first, c->np_last);
    } else if (c->precision == CAPAC_PREC_FP32_REFINED) {
        RelayG32 m{c->relay32, c->relay_ld, c->r, c->metal_u, c->bc};
        const int nm = c->bc.Nm;
        if (nm >= CAPAC_G32_LARGE_MIN) spmv_t<RelayG32, CAPAC_G32_ROWS_LARGE>(c, S, m, x, y, mode, part, sc);
        else if (nm >= CAPAC_G32_MID_MIN) spmv_t<RelayG32, CAPAC_G32_ROWS_MID>(c, S, m, x, y, mode, part, sc);
        else spmv_t<RelayG32, 1>(c, S, m, x, y, mode, part, sc);
    } else {
        RelayM64 m{c->r, c->bc};
        spmv_t<RelayM64, 1>(c, S, m, x, y, mode, part, sc);
    }
}

// the TILES product's tile launch alone (the fused GMRES column forms the
// sums itself); returns the input the sums take (x, or the corrected input
// with a coarse correction) and whether the sums form the Lagrange rows
static const double* spmv_tiles_launch(capac_ctx* c, const Solver* S, const double* x, double* y,
                                       const SolveScalars* sc, bool coarse, int* lag_in_sums) {
    const int nb = c->relay_nb, nt = (int)tile_count(nb);
    CoarseArgs cg{NULL, NULL, NULL, NULL, 0};
    if (coarse && c->cg_n) cg = CoarseArgs{c->cg_A, c->cg_s, reinterpret_cast<const float4*>(c->cg_w), S->t,
                                           c->cg_n};
    const int lag_t = cg.nc ? 0 : S->n_np;
    k_spmv_tiles<<<nt + lag_t, kTileThreads, 0, c->stream>>>(
        reinterpret_cast<const float4*>(c->relay32), nb, nt, S->Nm, S->n_np, c->np_of_metal, c->np_first,
        c->np_last, c->r, c->metal_u, c->bc, x, y, c->relay_part, c->relay_cnt, 0, S->r0, S->K, NULL, sc, cg);
    *lag_in_sums = cg.nc ? 1 : 0;
    return cg.nc ? cg.x2 : x;
}

struct SolveOut {
    int iters, kind;
    double rTr, bTb;
};

// BiCGSTAB on the resident relay (format by precision) for A x = b, x the
// warm start and the result.  With check64, every batch ends with the FP64
// residual b - A x (guarded: it runs only once the solve has finished), so
// the refinement test needs no extra synchronisation; (rr, rr) lands in
// res_host[0].
static int bicgstab(capac_ctx* c, Solver* S, const double* b, double* x, int first_batch, bool check64,
                    SolveOut* out) {
    const int n = S->n_mat, n4 = 4 * S->Nm;
    const int vb = S->n_vb;
    hipStream_t st = c->stream;
    spmv(c, S, x, S->Ax, 0, NULL, NULL);
    // the partial sums of the start go where the iterations' (r0, s) and
    // (r, r) partials go later; those are written before they are read
    k_bicg_init<<<vb, kVecThreads, 0, st>>>(n, n4, b, S->Ax, c->diag, S->r0, S->r, S->p, S->v, S->K, S->y,
                                            S->part2, S->part4);
    k_bicg_scalars<<<1, 256, 0, st>>>(vb, S->part2, S->part4, S->sc, c->solve_crit);
    // K1, K2 of iteration 1, then units (K3, K4 of n; K1, K2 of n+1): each
    // unit ends with a stop test, so a batch of units can end the solve
    spmv(c, S, S->y, S->v, 1, S->part1, S->sc);
    k_vec_s<<<vb, kVecThreads, 0, st>>>(n, 1, S->n_groups, S->part1, vb, S->part4, S->r, S->v, S->K,
                                        S->r0, S->s, S->z, S->part2, S->sc);
    // later batches start small: a first batch that falls short usually
    // needs only a few more iterations, and every unit launched after the
    // stop is a no-op that still costs four launches
    int it = 1, batch = first_batch > 1 ? first_batch : 2, launched = 0, next = 2;
    for (;;) {
        for (int u = 0; u < batch; ++u, ++it) {
            spmv(c, S, S->z, S->t, 2, S->part3, S->sc);
            k_vec_xp<<<vb, kVecThreads, 0, st>>>(n, it, S->n_groups, S->part3, vb, S->part2, x, S->y,
                                                 S->z, S->s, S->t, S->r, S->p, S->v, S->K, S->part4,
                                                 S->sc);
            spmv(c, S, S->y, S->v, 1, S->part1, S->sc);
            k_vec_s<<<vb, kVecThreads, 0, st>>>(n, it + 1, S->n_groups, S->part1, vb, S->part4, S->r,
                                                S->v, S->K, S->r0, S->s, S->z, S->part2, S->sc);
        }
        launched += batch;
        if (check64) {
            int n_part = 0;
            int rc = cpim_apply64(c, x, b, S->rr, S->part_res, S->sc, &n_part);
            if (rc) return rc;
            k_res_finish<<<1, 256, 0, st>>>(S->part_res, n_part, S->res, S->sc);
        }
        CHECK_HIP(hipGetLastError());
        if (S->in_red) {
            // the scalars, the FP64 residual and the neighbour-row flags in one transfer
            CHECK_HIP(hipMemcpyAsync(c->red_host + kRedFlags, c->red + kRedFlags,
                                     (kRedWords - kRedFlags) * sizeof(double), hipMemcpyDeviceToHost, st));
        } else {
            CHECK_HIP(hipMemcpyAsync(S->sc_host, S->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, st));
            CHECK_HIP(hipMemcpyAsync(S->res_host, S->res, sizeof(double), hipMemcpyDeviceToHost, st));
        }
        CHECK_HIP(hipStreamSynchronize(st));
        ++c->timing.solve_syncs;
        if (S->sc_host->done) break;
        if (launched > 1010) {
            capac_set_error("BiCGSTAB did not report completion after %d iterations", launched);
            return CAPAC_ERR_SOLVER;
        }
        batch = next;
        next = next < 16 ? 2 * next : 16;
    }
    out->iters = S->sc_host->iter;
    out->kind = S->sc_host->kind;
    out->rTr = S->sc_host->rTr;
    out->bTb = S->sc_host->bTb;
    return 0;
}

// GMRES on the FP32 relay for A x = b (kernels above), x the warm start and
// the result; batches of columns with the done flag, one synchronisation per
// batch, and the FP64 residual of the finished solve in the same transfer
// (check64), as bicgstab does.  A full basis restarts from the updated x.
static int gmres(capac_ctx* c, Solver* S, const double* b, double* x, int first_batch, bool check64,
                 SolveOut* out) {
    const int n = S->n_mat, n4 = 4 * S->Nm;
    const int vb = S->n_vb;
    hipStream_t st = c->stream;
    bool first = true;
    int launched = 0;
    // block preconditioner: the update applies B^-1 itself (block-aligned
    // launch of ug workgroups); the coarse correction is added by the tiled
    // product, or by its own launch with the rows layout
    const bool bjf = S->bj != NULL;
    const bool cg = bjf && c->cg_n > 0;
    const bool cg_fused = cg && c->relay_layout == CAPAC_LAYOUT_TILES && !CAPAC_TILE_FUSED;
    const int ug = bj_groups(tile_blocks(S->Nm)) + 1;
    // block preconditioner on the TILES layout: a column is three launches
    // (tiles, sums with the dots, update) instead of four
    const int nbk = tile_blocks(S->Nm);
    const bool fused = bjf && c->relay_layout == CAPAC_LAYOUT_TILES && !CAPAC_TILE_FUSED &&
                       nbk + S->n_np <= S->ph_rows;
    for (;;) {
        spmv(c, S, x, S->Ax, 0, NULL, NULL);
        k_gm_init<<<vb, kVecThreads, 0, st>>>(n, n4, b, S->Ax, c->diag, S->Q, S->p, S->K, S->part2, S->part4);
        k_gm_start<<<1, 256, 0, st>>>(vb, S->part2, S->part4, S->sc, S->gm, c->solve_crit, first ? 1 : 0);
        if (bjf) {
            bj_apply(c, S->Q, S->p, false, S->sc, 1);
            if (cg && !cg_fused) cg_apply(c, S->p, S->sc, 1);
        }
        first = false;
        // later batches start small (see bicgstab)
        int j = 0, batch = first_batch > 1 ? first_batch : 2, next = 2;
        for (;;) {
            for (int u = 0; u < batch && j <= kGmM; ++u, ++j) {
                // |q_j|^2 partials: the start's (vb) for q_0, the update's after
                const int nqq = j == 0 || !bjf ? vb : ug;
                int nph = vb;               // rows of part_h the update reduces
                if (j < kGmM && fused) {
                    // the product's tile launch, then its sums and the dots in one launch
                    int lag_here = 0;
                    const double* xin = spmv_tiles_launch(c, S, S->p, S->v, S->sc, cg_fused, &lag_here);
                    k_gm_tsum_dots<<<nbk + S->n_np + 1, kSumThreads, 0, st>>>(
                        nbk, S->Nm, S->n_np, n, j, c->np_of_metal, c->np_first, c->np_last, c->relay_part, c->bc,
                        xin, lag_here, S->v, S->Q, S->z, S->part4, nqq, S->gm, S->sc, S->part_h);
                    nph = nbk + S->n_np;
                } else {
                    if (j < kGmM) spmv(c, S, S->p, S->v, 0, NULL, S->sc, cg_fused);
                    const dim3 grid(j < kGmM ? vb : 1, j < kGmM ? (j + kGmDots) / kGmDots : 1);
                    k_gm_dots<<<grid, kVecThreads, 0, st>>>(n, j, S->v, S->Q, S->part4, nqq, S->gm, S->sc,
                                                            S->part_h, bjf ? S->z : NULL);
                }
                if (j < kGmM) {
                    if (bjf) {
                        k_gm_update_bj<<<ug, kVecThreads, 0, st>>>(
                            n, S->Nm, tile_blocks(S->Nm), j, S->v, S->z, S->Q, S->part_h, nph, S->gm, S->sc, S->p,
                            S->part4, S->bj, reinterpret_cast<const float4*>(c->cg_w), cg ? c->cg_s : NULL);
                        if (cg && !cg_fused) cg_apply(c, S->p, S->sc, 1);
                    } else {
                        k_gm_update<<<vb, kVecThreads, 0, st>>>(n, j, S->v, S->Q, S->K, S->part_h, vb, S->gm,
                                                                S->sc, S->p, S->part4);
                    }
                }
            }
            launched += batch;
            k_gm_finish<<<vb, kVecThreads, 0, st>>>(n, S->Q, S->K, S->gm, S->sc, x, bjf ? S->s : NULL);
            if (bjf) {
                bj_apply(c, S->s, x, true, S->sc, 2);
                cg_apply(c, x, S->sc, 2);
            }
            if (check64) {
                int n_part = 0;
                int rc = cpim_apply64(c, x, b, S->rr, S->part_res, S->sc, &n_part);
                if (rc) return rc;
                k_res_finish<<<1, 256, 0, st>>>(S->part_res, n_part, S->res, S->sc);
            }
            CHECK_HIP(hipGetLastError());
            if (S->in_red) {
                CHECK_HIP(hipMemcpyAsync(c->red_host + kRedFlags, c->red + kRedFlags,
                                         (kRedWords - kRedFlags) * sizeof(double), hipMemcpyDeviceToHost, st));
            } else {
                CHECK_HIP(hipMemcpyAsync(S->sc_host, S->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, st));
                CHECK_HIP(hipMemcpyAsync(S->res_host, S->res, sizeof(double), hipMemcpyDeviceToHost, st));
            }
            CHECK_HIP(hipStreamSynchronize(st));
            ++c->timing.solve_syncs;
            if (S->sc_host->done) break;
            if (j > kGmM || launched > 1100) {
                capac_set_error("GMRES did not report completion after %d columns", launched);
                return CAPAC_ERR_SOLVER;
            }
            batch = next;
            next = next < 16 ? 2 * next : 16;
        }
        if (S->sc_host->kind != kSolveRestart) break;
    }
    out->iters = S->sc_host->iter;
    out->kind = S->sc_host->kind;
    out->rTr = S->sc_host->rTr;
    out->bTb = S->sc_host->bTb;
    return 0;
}

// the inner solver: the reference's BiCGSTAB on the FP64 relay, GMRES on the FP32 one
static int inner_solve(capac_ctx* c, Solver* S, const double* b, double* x, int first_batch, bool check64,
                       SolveOut* out) {
    if (c->precision == CAPAC_PREC_FP32_REFINED) return gmres(c, S, b, x, first_batch, check64, out);
    return bicgstab(c, S, b, x, first_batch, check64, out);
}

// FP64 residual rr = b - A64 x with the relay recomputed on the fly; returns (rr, rr)
static int residual64(capac_ctx* c, Solver* S, const double* b, const double* x, double* rTr) {
    int n_part = 0;
    int rc = cpim_apply64(c, x, b, S->rr, S->part_res, NULL, &n_part);
    if (rc) return rc;
    k_res_finish<<<1, 256, 0, c->stream>>>(S->part_res, n_part, S->res, NULL);
    CHECK_HIP(hipGetLastError());
    CHECK_HIP(hipMemcpyAsync(S->res_host, S->res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    ++c->timing.solve_syncs;
    *rTr = S->res_host[0];
    return 0;
}

static void record(capac_ctx* c, const SolveOut& o) {
    c->solve_iters = o.iters;
    c->solve_kind = o.kind;
    c->solve_rTr = o.rTr;
    c->solve_bTb = o.bTb;
    c->timing.solve_iters += o.iters;
}

// Solves relay * x = b with the work vectors of S; x is the warm start and
// the result.  Fills c->solve_*; returns CAPAC_ERR_SOLVER when the reference
// would exit (bicg_stab.c:563-569).
static int solver_run_with(capac_ctx* c, Solver* S, const double* b, double* x) {
    // first batch: from the previous solves' iteration counts (warm starts
    // change slowly), so few launches are wasted as no-ops; the main solves'
    // counts alone (a refinement's correction solve has its own count, and
    // adding it would launch that many no-op units).  In MD with the
    // predictor the count swings from step to step, so the larger of the last
    // two counts, plus one, sizes the batch: a no-op column costs three short
    // launches, a batch that ends short a synchronisation and the batch-end
    // launches.  Au2057 FP32-refined, 300 MD steps (ms per step /
    // synchronisations per solve): mean of the last two plus one 1.231 /
    // 1.60, plus two 1.228 / 1.35, plus three 1.225 / 1.14, larger plus one
    // 1.223 / 1.18.
    int first = 2;
    if (c->batch_main > 1)
        first = (c->batch_main2 > c->batch_main ? c->batch_main2 : c->batch_main) + 1;
    const bool refine = c->precision == CAPAC_PREC_FP32_REFINED;
    // (the block-aligned update takes the Lagrange rows in one workgroup)
    S->bj = refine && c->bj_inv && c->bj_valid && S->n_np <= kVecThreads ? c->bj_inv : NULL;
    SolveOut o;
    int rc = inner_solve(c, S, b, x, first, refine, &o);
    if (rc) return rc;
    c->batch_main2 = c->batch_main;
    c->batch_main = o.iters;
    c->solve_refine_rounds = 0;
    if (refine && (o.kind == CAPAC_SOLVE_CONVERGED || o.kind == CAPAC_SOLVE_LOOSENED)) {
        // FP64 refinement: the reference's criterion on the true residual,
        // already computed at the end of the last batch
        const double crit = (o.kind == CAPAC_SOLVE_LOOSENED ? 100.0 : 1.0) * c->solve_crit * o.bTb;
        double rTr = S->res_host[0];
        int total = o.iters, rounds = 0;
        while (!(rTr < crit) && rounds < 4) {
            // the correction solves A d = rr / |rr| and x += |rr| d: with the
            // residual scaled to unit norm, the solver's absolute early exit
            // (bTb < 1e-10, meant for the system itself) cannot end a
            // correction whose residual is merely small
            const int n = S->n_mat;
            const double norm = sqrt(rTr);
            k_zero<<<(n + 255) / 256, 256, 0, c->stream>>>(n, S->d);
            k_scale<<<(n + 255) / 256, 256, 0, c->stream>>>(n, 1.0 / norm, S->rr);
            SolveOut od;
            const int first_corr = c->batch_corr > 0 ? c->batch_corr + 1 : 4;
            if ((rc = inner_solve(c, S, S->rr, S->d, first_corr, false, &od))) return rc;
            c->batch_corr = od.iters;
            total += od.iters;
            k_axpy<<<(n + 255) / 256, 256, 0, c->stream>>>(n, norm, S->d, x);
            if ((rc = residual64(c, S, b, x, &rTr))) return rc;
            ++rounds;
        }
        o.iters = total;
        o.rTr = rTr;
        if (!(rTr < crit)) o.kind = CAPAC_SOLVE_FAILED;
        c->solve_refine_rounds = rounds;
    }
    record(c, o);
    if (c->solve_kind == CAPAC_SOLVE_FAILED) {
        capac_set_error("Error: bicg_stab not converged. rTr= %e, bTb= %e", c->solve_rTr, c->solve_bTb);
        return CAPAC_ERR_SOLVER;
    }
    return 0;
}

int solver_run(capac_ctx* c, const double* b, double* x) { return solver_run_with(c, c->solver, b, x); }

// ---- the solve without library-allocated work space ----
// The work vectors, partial sums, solver scalars and residual of one solve
// live in a file-scope device array instead of memory the context allocated,
// so solver_run_noalloc calls no allocator.  Each device holds its own copy
// of the array at its own address, looked up once per device at context
// creation (solver_noalloc_init).  Every call on a device shares its array:
// calls that overlap on one device race.
constexpr int kNoAllocMaxUnknowns = 16384;
constexpr size_t kNoAllocGroups = kNoAllocMaxUnknowns + 1;       // metal row-groups + Lagrange rows, bound
constexpr size_t kNoAllocVecBlocks = (kNoAllocMaxUnknowns + kVecThreads - 1) / kVecThreads + 1;
// part4 also holds the block-aligned update's partials: at most 4096 metal atoms, 64 blocks
constexpr size_t kNoAllocPart4 = std::max(kNoAllocVecBlocks, (size_t)bj_groups(kNoAllocMaxUnknowns / 4 / kTileEdge) + 2);
constexpr size_t kNoAllocScWords = 16;                             // SolveScalars, padded
constexpr size_t kNoAllocGmWords = (sizeof(GmState) + 255) / 256 * 32;   // GmState, padded to 256 B
constexpr size_t kNoAllocWords = 12 * (size_t)kNoAllocMaxUnknowns + 5 * kNoAllocGroups +
                                 kNoAllocVecBlocks + kNoAllocPart4 + kNoAllocScWords + 4 +
                                 (size_t)(kGmM + 1) * kNoAllocMaxUnknowns + kNoAllocVecBlocks * (kGmM + 1) +
                                 kNoAllocGmWords;
static_assert(sizeof(SolveScalars) <= kNoAllocScWords * sizeof(double), "solver scalars overflow their slot");
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocWords];

static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// the device address of `scratch` on device dev (the current device)
int solver_noalloc_init(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        capac_set_error("device %d is beyond the %d devices the allocation-free solve supports", dev, kMaxDevices);
        return CAPAC_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void* p = NULL;
    CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
    return 0;
}

int solver_run_noalloc(capac_ctx* c, const double* b, double* x) {
    const int n = c->n_mat;
    if (n > kNoAllocMaxUnknowns) {
        capac_set_error("the allocation-free solve takes at most %d CPIM unknowns (this system has %d)",
                        kNoAllocMaxUnknowns, n);
        return CAPAC_ERR_ARG;
    }
    double* base;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : NULL;
    }
    if (!base) {
        capac_set_error("allocation-free work space of device %d was not set up", c->device);
        return CAPAC_ERR_HIP;
    }
    const Solver* own = c->solver;
    Solver V;
    memset(&V, 0, sizeof V);
    V.n_mat = n; V.Nm = own->Nm; V.n_np = own->n_np; V.n_groups = own->n_groups; V.n_vb = own->n_vb;
    double* p = base;
    double** vecs[12] = {&V.r0, &V.r, &V.p, &V.v, &V.s, &V.t, &V.y, &V.z, &V.K, &V.Ax, &V.d, &V.rr};
    for (int k = 0; k < 12; ++k) { *vecs[k] = p; p += kNoAllocMaxUnknowns; }
    V.part1 = p; p += kNoAllocGroups;
    V.part3 = p; p += 3 * kNoAllocGroups;
    V.part_res = p; p += kNoAllocGroups;
    V.part2 = p; p += kNoAllocVecBlocks;
    V.part4 = p; p += kNoAllocPart4;
    V.sc = (SolveScalars*)p; p += kNoAllocScWords;
    V.res = p; p += 4;
    V.Q = p; p += (size_t)(kGmM + 1) * kNoAllocMaxUnknowns;
    V.part_h = p; p += kNoAllocVecBlocks * (kGmM + 1);
    V.ph_rows = (int)kNoAllocVecBlocks;
    V.gm = (GmState*)p; p += kNoAllocGmWords;
    V.sc_host = own->sc_host;
    V.res_host = own->res_host;
    V.in_red = false;
    // the array keeps the previous call's contents: clear it as a new
    // context's work space starts (a fill kernel, no memset node)
    k_zero<<<(unsigned)((kNoAllocWords + 255) / 256), 256, 0, c->stream>>>((int)kNoAllocWords, base);
    CHECK_HIP(hipGetLastError());
    return solver_run_with(c, &V, b, x);
}

// Warm start of an MD step's solve.  The reference starts from the previous
// step's solution (bicg_stab.c:303-306, order 0).  Along a trajectory the
// solution changes smoothly, so extrapolating the last two or three
// solutions (order 1, 2) starts much closer: on the Au561 deck the
// iterations to the same stop test fall from ~38 to ~23 (linear) and ~16
// (quadratic).  The stop test, and so the accuracy of every solution, is
// unchanged.  With fewer stored solutions than the order needs, the order
// drops to what is there.
// The extrapolation runs in the next force call's prologue (force_prologue).
void solver_predict(capac_ctx* c) {
    const int order = c->pred_order < c->pred_count - 1 ? c->pred_order : c->pred_count - 1;
    c->pred_pending = (order > 0 && c->n_mat > 0) ? order : 0;
}

// clears red[0..n_red), applies a pending extrapolation and saves x in x_backup
int force_prologue(capac_ctx* c, int n_red) {
    const int n = c->n_mat;
    const int work = n > n_red ? n : n_red;
    k_force_prologue<<<(work + 255) / 256, 256, 0, c->stream>>>(n, c->pred_pending, c->x_hist[0], c->x_hist[1],
                                                                c->x_hist[2], c->x, c->x_backup, c->red, n_red);
    c->pred_pending = 0;
    CHECK_HIP(hipGetLastError());
    return 0;
}

void solver_push_history(capac_ctx* c) {
    if (c->n_mat == 0) return;
    double* oldest = c->x_hist[2];
    c->x_hist[2] = c->x_hist[1];
    c->x_hist[1] = c->x_hist[0];
    c->x_hist[0] = oldest;
    // (the force call's CPIM force launch has copied the solution already
    // when it was asked to, into this same buffer)
    if (c->hist_copied != oldest)
        (void)hipMemcpyAsync(oldest, c->x, sizeof(double) * c->n_mat, hipMemcpyDeviceToDevice, c->stream);
    c->hist_dst = NULL;
    c->hist_copied = NULL;
    if (c->pred_count < 3) ++c->pred_count;
}

// a solver work vector that kernel timing may overwrite (rebuilt by every solve)
double* solver_scratch(capac_ctx* c) { return c->solver->Ax; }

// one plain SpMV over the resident matrix (for kernel timing)
int solver_spmv_only(capac_ctx* c, const double* x, double* y) {
    spmv(c, c->solver, x, y, 0, NULL, NULL);
    CHECK_HIP(hipGetLastError());
    return 0;
}
