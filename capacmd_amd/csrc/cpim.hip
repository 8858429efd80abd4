// Capacitance-polarizability interaction model (CPIM) on the device:
//   cpim_rhs       mpi_cpff_vec_ext        (cpff.c:43-198)
//   cpim_assemble  mpi_cpff_mat_relay_COO  (cpff.c:205-519)
//   cpim_forces    mpi_cpff_force          (cpff.c:526-950)
//
// The relay has no cut-off (every metal pair), and its blocks are given in
// relay.h.  The FP64 path stores no matrix: its SpMV recomputes each block
// from the positions (the entries the reference assembles, term by term).
// The FP32-refined path stores four FP32 coefficients per ordered pair,
// component-major (plane k of pair (i, j) at relay32[(i*4 + k)*ld + j], ld =
// Nm rounded up to 64), so lanes that take consecutive j read and write 256
// contiguous bytes per instruction.  The Lagrange rows and columns of the
// charge-neutrality constraint are applied analytically by the solver.
#include "capac_internal.h"
#include "erfc_fast.h"
#include "relay.h"
#include "relay_tile.h"
#include "sc_terms.h"

namespace {

// ------------------------------------------------------------------ RHS
// One workgroup per metal atom over the non-metal front of its row (four
// wavefronts split the row, as in the metal-side force kernels).
// (metal atom i's rows of b; the caller has staged the erfcx table)
__device__ __forceinline__ void rhs_block(const double* __restrict__ r, const double* __restrict__ charge,
                                          const BoxConst& bc, const int* __restrict__ n_front,
                                          const int* __restrict__ nbr, int stride,
                                          const double* __restrict__ np_charge, double* __restrict__ b, int i,
                                          const double* s_erfcx, double (*s_w)[4]) {
    const int N = bc.N, Nm = bc.Nm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
    int nf = n_front[i];
    const int* row = nbr + (size_t)i * stride;
    double ex = 0.0, ey = 0.0, ez = 0.0, pm = 0.0;
    for (int k = threadIdx.x; k < nf; k += blockDim.x) {
        int a = row[k];
        double q = charge[a];
        if (q == 0.0) continue;
        // rvec = r_metal - r_atom (cpff.c:105-108)
        double dx = xi - r[a], dy = yi - r[N + a], dz = zi - r[2 * N + a];
        pbc12(bc, dx, dy, dz);
        double r2 = dist2_exact(dx, dy, dz);
        if (!(r2 < bc.rc2)) continue;             // rows reach rc + skin
        double rij = sqrt(r2);
        double e_field, v_field;
        if (bc.use_coulomb == 1) {
            double e2 = exp_poly<false>(-bc.alpha * bc.alpha * r2);
            double erfc_a = erfc_from_exp(bc.alpha * rij, e2, s_erfcx);
            e_field = K_FQQ * q * (erfc_a / r2 + bc.a2_sqrtPI * e2 / rij - bc.wolfConst) / rij;
            v_field = K_FQQ * q * (erfc_a / rij - bc.erfc_arCut + bc.wolfConst * (rij - bc.rc));
        } else {
            e_field = K_FQQ * q / (r2 * rij);
            v_field = K_FQQ * q / rij;
        }
        double s = e_field * K_CONV_E;
        ex += s * dx; ey += s * dy; ez += s * dz;
        pm += v_field * K_CONV_V;
    }
    ex = wave_sum(ex); ey = wave_sum(ey); ez = wave_sum(ez); pm = wave_sum(pm);
    if (lane == 0) { s_w[wave][0] = ex; s_w[wave][1] = ey; s_w[wave][2] = ez; s_w[wave][3] = pm; }
    __syncthreads();
    if (threadIdx.x == 0) {
        ex = s_w[0][0] + s_w[1][0] + s_w[2][0] + s_w[3][0];
        ey = s_w[0][1] + s_w[1][1] + s_w[2][1] + s_w[3][1];
        ez = s_w[0][2] + s_w[1][2] + s_w[2][2] + s_w[3][2];
        pm = s_w[0][3] + s_w[1][3] + s_w[2][3] + s_w[3][3];
        b[3 * i + 0] = ex + bc.efield[0] * K_CONV_E;
        b[3 * i + 1] = ey + bc.efield[1] * K_CONV_E;
        b[3 * i + 2] = ez + bc.efield[2] * K_CONV_E;
        double pot_ext = bc.efield[0] * xi + bc.efield[1] * yi + bc.efield[2] * zi;
        b[3 * Nm + i] = pm - pot_ext * K_CONV_V;
        if (i < bc.n_np) b[4 * Nm + i] = np_charge[i];
    }
}

__global__ void __launch_bounds__(256)
k_rhs(const double* __restrict__ r, const double* __restrict__ charge, BoxConst bc,
      const int* __restrict__ n_front, const int* __restrict__ nbr, int stride,
      const double* __restrict__ np_charge, double* __restrict__ b) {
    main_stream_priority();
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    __shared__ double s_w[4][4];
    load_erfcx_table(s_erfcx);
    __syncthreads();
    if ((int)blockIdx.x >= bc.Nm) return;
    rhs_block(r, charge, bc, n_front, nbr, stride, np_charge, b, blockIdx.x, s_erfcx, s_w);
}

// ------------------------------------------------------------------ relay
// FP32-refined relay: the four coefficient planes of relay.h, T-tensors in
// FP32 from the separation formed in FP64 with the reference's pbc_12 and
// rounded once to FP32.  The diagonal pairs hold zeros (their blocks are
// applied analytically).  One thread per ordered pair: grid (ceil(Nm/256), Nm).
__global__ void __launch_bounds__(256)
k_assemble32(const double* __restrict__ r, BoxConst bc, int ld, float* __restrict__ relay) {
    main_stream_priority();
    const int Nm = bc.Nm;
    int i = blockIdx.y;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Nm) return;
    float cf[kRelay32Planes] = {0.f, 0.f, 0.f, 0.f};
    if (i != j) {
        double dx, dy, dz;
        metal_sep(r, bc, i, j, dx, dy, dz);
        float rx = (float)dx, ry = (float)dy, rz = (float)dz;
        float r2 = (rx * rx + ry * ry) + rz * rz;
        TCoef32 t = t_coef32(r2, (float)bc.inv_R_pp, (float)bc.inv_R_pq, (float)bc.inv_R_qq);
        cf[0] = t.c2 - 3.0f * t.c1;
        cf[1] = t.c1 * r2;
        cf[2] = t.c3;
        cf[3] = t.cq;
    }
    float* dst = relay + (size_t)i * kRelay32Planes * ld + j;
#pragma unroll
    for (int k = 0; k < kRelay32Planes; ++k) dst[(size_t)k * ld] = cf[k];
}

// The same coefficients in the TILES layout (relay_tile.h): one workgroup
// per tile, the threads over its 4096 slots; each unordered pair is formed
// once, from metal_sep(i, j) with i in the tile's row block (the reversed
// separation is its exact negative, and the coefficients depend on r^2
// only).  Slot k = 0 of a diagonal tile (i = j) and pairs past Nm hold zeros.
__device__ __forceinline__ void assemble_tile_block(const double* __restrict__ r, const BoxConst& bc, int Nb,
                                                    float4* __restrict__ tiles, int t) {
    const int Nm = bc.Nm;
    int I, J;
    tile_coords(t, Nb, I, J);
    float4* dst = tiles + (size_t)t * kTileSlots;
    for (int e = threadIdx.x; e < kTileSlots; e += blockDim.x) {
        const int k = e / kTileEdge, l = e % kTileEdge;
        const int i = I * kTileEdge + l, j = J * kTileEdge + ((l + k) & (kTileEdge - 1));
        float4 cf = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i != j && i < Nm && j < Nm) {
            double dx, dy, dz;
            metal_sep(r, bc, i, j, dx, dy, dz);
            float rx = (float)dx, ry = (float)dy, rz = (float)dz;
            float r2 = (rx * rx + ry * ry) + rz * rz;
            TCoef32 t = t_coef32(r2, (float)bc.inv_R_pp, (float)bc.inv_R_pq, (float)bc.inv_R_qq);
            cf = make_float4(t.c2 - 3.0f * t.c1, t.c1 * r2, t.c3, t.cq);
        }
        dst[e] = cf;
    }
}

__global__ void __launch_bounds__(256)
k_assemble32_tiles(const double* __restrict__ r, BoxConst bc, int Nb, float4* __restrict__ tiles) {
    main_stream_priority();
    assemble_tile_block(r, bc, Nb, tiles, blockIdx.x);
}

// Metal separations from metal atom 0, d_j = pbc_12(r_j - r_0) in bohr, and
// whether they are compact: when every component's spread max_j d_j -
// min_j d_j is below half the box edge, every pair's minimum-image
// separation is d_j - d_i (pbc_12 leaves a difference inside half the box
// alone), and the FP32 product applies its blocks from d without the
// per-pair image test.  One workgroup; the flag lands at u[4 Nm].
__device__ __forceinline__ void metal_unwrap_block(const double* __restrict__ r, const BoxConst& bc,
                                                   double* __restrict__ u, double (*s_lo)[16], double (*s_hi)[16]) {
    const int N = bc.N, Nm = bc.Nm;
    const double x0 = r[0], y0 = r[N], z0 = r[2 * N];
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < Nm; j += blockDim.x) {
        double d[3] = {r[j] - x0, r[N + j] - y0, r[2 * N + j] - z0};
        pbc12(bc, d[0], d[1], d[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], d[k]);
            hi[k] = fmax(hi[k], d[k]);
            u[4 * j + k] = d[k] * K_NM2BOHR;
        }
        u[4 * j + 3] = 0.0;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], off, 64));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], off, 64));
        }
    if (lane == 0)
        for (int k = 0; k < 3; ++k) { s_lo[k][wave] = lo[k]; s_hi[k][wave] = hi[k]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool compact = true;
        for (int k = 0; k < 3; ++k) {
            double a = s_lo[k][0], b = s_hi[k][0];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { a = fmin(a, s_lo[k][w]); b = fmax(b, s_hi[k][w]); }
            compact = compact && (b - a) < bc.half[k] * (1.0 - 1e-9);
        }
        u[4 * Nm] = compact ? 1.0 : 0.0;
    }
}

__global__ void __launch_bounds__(1024)
k_metal_unwrap(const double* __restrict__ r, BoxConst bc, double* __restrict__ u) {
    main_stream_priority();
    __shared__ double s_lo[3][16], s_hi[3][16];
    metal_unwrap_block(r, bc, u, s_lo, s_hi);
}

// The right-hand side and the FP32 relay's TILES layout in one launch
// (they read the positions only and write different arrays): workgroup 0
// the metal unwrap, then the n_tiles relay tiles (the longest-running
// workgroups, so they start early), then one per metal atom for b, then the
// Sutton-Chen densities (N_m / 4 workgroups, when the grid has them).
__global__ void __launch_bounds__(256)
k_matrix_terms(const double* __restrict__ r, const double* __restrict__ charge, BoxConst bc,
               const int* __restrict__ n_front, const int* __restrict__ nbr, int stride,
               const double* __restrict__ np_charge, double* __restrict__ b, int Nb, int n_tiles,
               float4* __restrict__ tiles, double* __restrict__ u, const int* __restrict__ n_back,
               double* __restrict__ isd) {
    main_stream_priority();
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    __shared__ double s_w[4][4];
    __shared__ double s_lo[3][16], s_hi[3][16];
    const int blk = blockIdx.x;
    if (blk == 0) {
        metal_unwrap_block(r, bc, u, s_lo, s_hi);
    } else if (blk <= n_tiles) {
        assemble_tile_block(r, bc, Nb, tiles, blk - 1);
    } else if (blk > n_tiles + bc.Nm) {
        // the Sutton-Chen densities, four metal atoms per workgroup
        sc_dens_block(blk - 1 - n_tiles - bc.Nm, r, bc, n_back, nbr, stride, isd);
    } else {
        load_erfcx_table(s_erfcx);
        __syncthreads();
        rhs_block(r, charge, bc, n_front, nbr, stride, np_charge, b, blk - 1 - n_tiles, s_erfcx, s_w);
    }
}

// y = A x (or b - A x) with the FP64 relay recomputed pair by pair
// (relay_block64); one workgroup per row-group (metal i, then the Lagrange
// rows); part[g] = (out, out) over the group's rows
__global__ void __launch_bounds__(256)
k_apply64(const double* __restrict__ r, BoxConst bc, const int* __restrict__ np_of,
          const int* __restrict__ np_first, const int* __restrict__ np_last,
          const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ out,
          double* __restrict__ part, const SolveScalars* __restrict__ guard) {
    main_stream_priority();
    __shared__ double s_w[4][4];
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    if (guard && !(guard->done && (guard->kind == CAPAC_SOLVE_CONVERGED || guard->kind == CAPAC_SOLVE_LOOSENED)))
        return;
    load_erfcx_table(s_erfcx);
    __syncthreads();
    const int Nm = bc.Nm, N = bc.N;
    const int g = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double o[4] = {0.0, 0.0, 0.0, 0.0};
    if (g >= Nm) {
        const int k = g - Nm;
        for (int mi = np_first[k] + threadIdx.x; mi <= np_last[k]; mi += blockDim.x) o[3] += x[3 * Nm + mi];
    } else {
        // block (i, j) as g_a r r^T + g_b 1, c3 r; -c3 r^T, cq (relay.h) from
        // the FP64 coefficients of t_coef_fast, applied in that form
        const int i = g;
        const double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
        for (int j = threadIdx.x; j < Nm; j += blockDim.x) {
            const double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2], x3 = x[3 * Nm + j];
            if (j == i) {
                o[0] += bc.inv_polar * x0;
                o[1] += bc.inv_polar * x1;
                o[2] += bc.inv_polar * x2;
                o[3] += -bc.inv_capac * x3;
                continue;
            }
            double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
            pbc12(bc, dx, dy, dz);
            dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
            const double r2 = dist2_exact(dx, dy, dz);
            const TCoef t = t_coef_fast(bc, r2, s_erfcx);
            block32_apply(dx, dy, dz, t.c2 - 3.0 * t.c1, t.c1 * r2, t.c3, t.cq, x0, x1, x2, x3, o);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = wave_sum(o[c]);
    if (lane == 0)
        for (int c = 0; c < 4; ++c) s_w[wave][c] = o[c];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int c = 0; c < 4; ++c) o[c] = s_w[0][c] + s_w[1][c] + s_w[2][c] + s_w[3][c];
    double acc = 0.0;
    if (g >= Nm) {
        const int row = 4 * Nm + (g - Nm);
        double val = b ? b[row] - o[3] : o[3];
        out[row] = val;
        acc = val * val;
    } else {
        const int i = g;
        const int np = np_of[i];
        if (np >= 0) o[3] += x[4 * Nm + np];
        const int rows[4] = {3 * i, 3 * i + 1, 3 * i + 2, 3 * Nm + i};
        for (int c = 0; c < 4; ++c) {
            double val = b ? b[rows[c]] - o[c] : o[c];
            out[rows[c]] = val;
            acc += val * val;
        }
    }
    if (part) part[g] = acc;
}

// The same product on the TILES layout's tiles, each unordered metal pair
// once: one workgroup per tile (I, J), four wavefronts of sixteen diagonals;
// lane l applies block (i, j) to x_j for its row i (registers) and block
// (j, i) -- the same coefficients at separation -r -- to x_i for row j (the
// wavefront's column sums in LDS; the 64 lanes of a step meet 64 different
// j).  A diagonal tile takes every ordered pair of its block for the rows
// only (k = 1..63), as the FP32 product does.  The four wavefronts' sums go
// to the partials P[I][J] / P[J][I] (4 x 64 doubles, the FP32 product's
// layout), which k_apply64_sum adds.
__global__ void __launch_bounds__(256)
k_apply64_tiles(const double* __restrict__ r, BoxConst bc, const double* __restrict__ x, int Nb,
                double* __restrict__ P, const SolveScalars* __restrict__ guard) {
    main_stream_priority();
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    __shared__ double s_j[7][kTileEdge];            // block J: positions (nm) and x (4)
    __shared__ double s_col[4][4][kTileEdge];
    __shared__ double s_row[4][4][kTileEdge];
    if (guard && !(guard->done && (guard->kind == CAPAC_SOLVE_CONVERGED || guard->kind == CAPAC_SOLVE_LOOSENED)))
        return;
    const int N = bc.N, Nm = bc.Nm;
    int I, J;
    tile_coords(blockIdx.x, Nb, I, J);
    load_erfcx_table(s_erfcx);
    if (threadIdx.x < kTileEdge) {
        const int j = J * kTileEdge + threadIdx.x;
        const bool ok = j < Nm;
        s_j[0][threadIdx.x] = ok ? r[j] : 0.0;
        s_j[1][threadIdx.x] = ok ? r[N + j] : 0.0;
        s_j[2][threadIdx.x] = ok ? r[2 * N + j] : 0.0;
        s_j[3][threadIdx.x] = ok ? x[3 * j] : 0.0;
        s_j[4][threadIdx.x] = ok ? x[3 * j + 1] : 0.0;
        s_j[5][threadIdx.x] = ok ? x[3 * j + 2] : 0.0;
        s_j[6][threadIdx.x] = ok ? x[3 * Nm + j] : 0.0;
    }
    for (int k = threadIdx.x; k < 4 * 4 * kTileEdge; k += blockDim.x) (&s_col[0][0][0])[k] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = I * kTileEdge + lane;
    const bool iok = i < Nm;
    const double xi = iok ? r[i] : 0.0, yi = iok ? r[N + i] : 0.0, zi = iok ? r[2 * N + i] : 0.0;
    const double vi[4] = {iok ? x[3 * i] : 0.0, iok ? x[3 * i + 1] : 0.0, iok ? x[3 * i + 2] : 0.0,
                          iok ? x[3 * Nm + i] : 0.0};
    const bool diag = I == J;
    double orow[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int s = 0; s < 16; ++s) {
        const int k = wave * 16 + s;
        const int jl = (lane + k) & (kTileEdge - 1);
        if (!iok || J * kTileEdge + jl >= Nm || (diag && k == 0)) continue;
        double dx = s_j[0][jl] - xi, dy = s_j[1][jl] - yi, dz = s_j[2][jl] - zi;
        pbc12(bc, dx, dy, dz);
        dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
        const double r2 = dist2_exact(dx, dy, dz);
        const TCoef t = t_coef_fast(bc, r2, s_erfcx);
        const double ga = t.c2 - 3.0 * t.c1, gb = t.c1 * r2;
        block32_apply(dx, dy, dz, ga, gb, t.c3, t.cq, s_j[3][jl], s_j[4][jl], s_j[5][jl], s_j[6][jl], orow);
        if (!diag) {
            double oc[4] = {0.0, 0.0, 0.0, 0.0};
            block32_apply(-dx, -dy, -dz, ga, gb, t.c3, t.cq, vi[0], vi[1], vi[2], vi[3], oc);
#pragma unroll
            for (int c = 0; c < 4; ++c) s_col[wave][c][jl] += oc[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) s_row[wave][c][lane] = orow[c];
    __syncthreads();
    const int c = threadIdx.x >> 6, l = threadIdx.x & 63;
    double rs = 0.0, cs = 0.0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        rs += s_row[w][c][l];
        cs += s_col[w][c][l];
    }
    const size_t blk = 4 * kTileEdge;
    P[((size_t)I * Nb + J) * blk + threadIdx.x] = rs;
    if (!diag) P[((size_t)J * Nb + I) * blk + threadIdx.x] = cs;
}

// The rows' outputs from the tiles' partials: workgroup B < Nb the 4 x 64
// rows of metal block B (thread (c, l): the Nb partials in order, the
// diagonal term, the Lagrange column), then one workgroup per nanoparticle's
// Lagrange row; out = b - A x (b non-NULL) or A x, part[g] = (out, out) over
// workgroup g's rows.
__global__ void __launch_bounds__(256)
k_apply64_sum(BoxConst bc, const int* __restrict__ np_of, const int* __restrict__ np_first,
              const int* __restrict__ np_last, const double* __restrict__ x, const double* __restrict__ b, int Nb,
              const double* __restrict__ P, double* __restrict__ out, double* __restrict__ part,
              const SolveScalars* __restrict__ guard) {
    main_stream_priority();
    __shared__ double s_w[4];
    if (guard && !(guard->done && (guard->kind == CAPAC_SOLVE_CONVERGED || guard->kind == CAPAC_SOLVE_LOOSENED)))
        return;
    const int Nm = bc.Nm;
    const int g = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc = 0.0;
    if (g < Nb) {
        const int c = wave, i = g * kTileEdge + lane;
        double o = 0.0;
        for (int S0 = 0; S0 < Nb; S0 += 16) {
            double pv[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
                pv[q] = S0 + q < Nb ? P[((size_t)g * Nb + S0 + q) * (4 * kTileEdge) + threadIdx.x] : 0.0;
#pragma unroll
            for (int q = 0; q < 16; ++q) o += pv[q];
        }
        if (i < Nm) {
            if (c < 3) {
                o += bc.inv_polar * x[3 * i + c];
            } else {
                o += -bc.inv_capac * x[3 * Nm + i];
                const int np = np_of[i];
                if (np >= 0) o += x[4 * Nm + np];
            }
            const int row = c < 3 ? 3 * i + c : 3 * Nm + i;
            const double val = b ? b[row] - o : o;
            out[row] = val;
            acc = val * val;
        }
    } else {
        const int k = g - Nb;
        double o3 = 0.0;
        for (int mi = np_first[k] + (int)threadIdx.x; mi <= np_last[k]; mi += blockDim.x) o3 += x[3 * Nm + mi];
        o3 = wave_sum(o3);
        if (lane == 0) s_w[wave] = o3;
        __syncthreads();
        if (threadIdx.x == 0) {
            o3 = s_w[0] + s_w[1] + s_w[2] + s_w[3];
            const int row = 4 * Nm + k;
            const double val = b ? b[row] - o3 : o3;
            out[row] = val;
            if (part) part[g] = val * val;
        }
        return;
    }
    acc = wave_sum(acc);
    if (lane == 0) s_w[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0 && part) part[g] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ void k_diag(BoxConst bc, int n_mat, double* __restrict__ diag) {
    main_stream_priority();
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_mat) return;
    int Nm = bc.Nm;
    diag[k] = k < 3 * Nm ? bc.inv_polar : (k < 4 * Nm ? -bc.inv_capac : 1.0);
}

// dense 4Nm x 4Nm metal block of either relay (tests): FP32 from the stored
// coefficients and the separation, FP64 from the positions as the SpMV forms it
__global__ void k_dense(const float* __restrict__ relay, int tiles, int Nb, const double* __restrict__ r,
                        BoxConst bc, int ld, double* __restrict__ dense) {
    const int Nm = bc.Nm;
    long n4 = 4L * Nm;
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n4 * n4) return;
    long row = idx / n4, col = idx % n4;
    int i = row < 3L * Nm ? (int)(row / 3) : (int)(row - 3L * Nm);
    int a = row < 3L * Nm ? (int)(row % 3) : 3;
    int j = col < 3L * Nm ? (int)(col / 3) : (int)(col - 3L * Nm);
    int b = col < 3L * Nm ? (int)(col % 3) : 3;
    double v;
    if (i == j) {
        v = (a != b) ? 0.0 : (a < 3 ? bc.inv_polar : -bc.inv_capac);
    } else if (relay) {
        double ga, gb, c3, cq;
        if (tiles) {
            const float4 cf = reinterpret_cast<const float4*>(relay)[tile_slot(i, j, Nb)];
            ga = cf.x; gb = cf.y; c3 = cf.z; cq = cf.w;
        } else {
            const float* cp = relay + (size_t)i * kRelay32Planes * ld + j;
            ga = cp[0]; gb = cp[ld]; c3 = cp[2 * ld]; cq = cp[3 * ld];
        }
        double rv[3];
        metal_sep(r, bc, i, j, rv[0], rv[1], rv[2]);
        if (a < 3 && b < 3) v = ga * rv[a] * rv[b] + (a == b ? gb : 0.0);
        else if (a < 3) v = c3 * rv[a];
        else if (b < 3) v = -c3 * rv[b];
        else v = cq;
    } else {
        double blk[16];
        relay_block64(r, bc, i, j, blk);
        v = blk[a * 4 + b];
    }
    dense[idx] = v;
}

// ------------------------------------------------------------------ forces
// Metal-water terms (cpff.c:581-721), evaluated for r = r_metal - r_atom in nm:
// DSF charge(q_i)-charge(q_a) and induced dipole(p_i)-charge(q_a).
__device__ __forceinline__ void mw_pair(const BoxConst& bc, double dx, double dy, double dz, double r2,
                                        double qa, double qm, const double* pm, double f[3],
                                        double* e10, double* e11, const double* __restrict__ erfcx) {
    double rij = sqrt(r2);
    double rv[3] = {dx, dy, dz};
    double const_a0, erfc_a = 0.0, e2 = 0.0;
    if (bc.use_coulomb == 1) {
        // erfc from the Gaussian the term needs anyway (the erfcx table)
        e2 = exp_poly<false>(-bc.alpha * bc.alpha * r2);
        erfc_a = erfc_from_exp(bc.alpha * rij, e2, erfcx);
        const_a0 = (erfc_a / r2 + bc.a2_sqrtPI * e2 / rij - bc.wolfConst) / rij;
        *e10 += K_FQQ * qa * qm * (erfc_a / rij - bc.erfc_arCut + bc.wolfConst * (rij - bc.rc));
    } else {
        const_a0 = 1.0 / (r2 * rij);
        *e10 += K_FQQ * qa * qm / rij;
    }
    double sq = const_a0 * K_FQQ * qa * qm;
    f[0] = sq * dx; f[1] = sq * dy; f[2] = sq * dz;
    // dipole in e*nm
    double p[3] = {pm[0] / K_NM2BOHR, pm[1] / K_NM2BOHR, pm[2] / K_NM2BOHR};
    double pr = p[0] * dx + p[1] * dy + p[2] * dz;
    double A1, B1 = 0.0, C1 = 0.0;
    double inv_r = 1.0 / rij;
    double inv_r3 = inv_r * inv_r * inv_r;
    if (bc.use_coulomb == 1) {
        A1 = inv_r3 * inv_r * inv_r * (erfc_a + bc.a2_sqrtPI * rij * e2);
        double alpha3 = bc.alpha * bc.alpha * bc.alpha;
        B1 = 4.0 * K_INV_SQRT_PI * alpha3 / r2 * e2;
        C1 = bc.wolfConst * inv_r3;
    } else {
        A1 = inv_r3 * inv_r * inv_r;
    }
    // f_b += -sum_a [A1 (3 r_a r_b - d_ab r^2) + B1 r_a r_b - C1 (r_a r_b - d_ab r^2)] FQQ q_a p_a
    double s = K_FQQ * qa;
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
        double t = A1 * (3.0 * pr * rv[bb] - p[bb] * r2) + B1 * pr * rv[bb] - C1 * (pr * rv[bb] - p[bb] * r2);
        f[bb] -= t * s;
    }
    *e11 += -K_FQQ * qa * pr * const_a0;
}

// ---------------------------------------------------------- force terms
// metal-water terms of metal atom i over its charged non-metal neighbours,
// the row split over 256 threads (t = 0..255); vals = e10, e11, virial
__device__ __forceinline__ void mw_metal_terms(const double* __restrict__ r, const double* __restrict__ charge,
                                               const BoxConst& bc, const int* __restrict__ n_front,
                                               const int* __restrict__ nbr, int stride,
                                               const double* __restrict__ x, int i, int t,
                                               double& fx, double& fy, double& fz, double vals[13],
                                               const double* __restrict__ erfcx) {
    const int N = bc.N, Nm = bc.Nm;
    const double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
    const double qm = x[3 * Nm + i];
    const double pm[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const int nf = n_front[i];
    const int* row = nbr + (size_t)i * stride;
    double e10 = 0.0, e11 = 0.0, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = t; k < nf; k += 256) {
        int a = row[k];
        double qa = charge[a];
        if (qa == 0.0) continue;
        double dx = xi - r[a], dy = yi - r[N + a], dz = zi - r[2 * N + a];
        pbc12(bc, dx, dy, dz);
        double r2 = dist2_exact(dx, dy, dz);
        if (!(r2 < bc.rc2)) continue;         // rows reach rc + skin
        double fp[3];
        mw_pair(bc, dx, dy, dz, r2, qa, qm, pm, fp, &e10, &e11, erfcx);
        fx += fp[0]; fy += fp[1]; fz += fp[2];
        double rv[3] = {dx, dy, dz};
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
            for (int b2 = 0; b2 < 3; ++b2) v[a2 * 3 + b2] -= rv[a2] * fp[b2];
    }
    vals[0] = e10; vals[1] = e11;
#pragma unroll
    for (int c = 0; c < 9; ++c) vals[2 + c] = v[c];
}

// One metal-metal pair (cpff.c:762-927), both orders at once: r = r_j - r_i
// in bohr, (qi, pi) and (qj, pj) the induced charges and dipoles.  The force
// on i is -g and on j +g: swapping i and j negates r and every term of g is
// odd in r once the pair's roles swap too, so the reference's two ordered
// evaluations give -g and +g up to rounding.  e12, e13, e14 receive the
// pair's energies (the two ordered halves of the reference together).  Each
// erf comes from the Gaussian the same term needs anyway (erf_from_exp, with
// the erfcx table staged in LDS).
struct MMConst {
    double iRqq, iRpq, iRpp, iRpq3, iRpp3, iRpp5;
};

__device__ __forceinline__ MMConst mm_const(const BoxConst& bc) {
    MMConst k;
    k.iRqq = bc.inv_R_qq; k.iRpq = bc.inv_R_pq; k.iRpp = bc.inv_R_pp;
    k.iRpq3 = k.iRpq * k.iRpq * k.iRpq;
    k.iRpp3 = k.iRpp * k.iRpp * k.iRpp;
    k.iRpp5 = k.iRpp3 * k.iRpp * k.iRpp;
    return k;
}

__device__ __forceinline__ void mm_pair(const MMConst& k, double dx, double dy, double dz, double qi,
                                        const double pi[3], double qj, const double pj[3],
                                        const double* __restrict__ erfcx, double g[3], double& e12, double& e13,
                                        double& e14) {
    const double r2 = dist2_exact(dx, dy, dz);
    const double inv_r = rsqrt_pos(r2), inv_r2 = inv_r * inv_r;
    const double rij = r2 * inv_r;
    const double inv_r3 = inv_r2 * inv_r, inv_r5 = inv_r3 * inv_r2, inv_r7 = inv_r5 * inv_r2;
    const double rv[3] = {dx, dy, dz};
    // charge-charge, T0
    const double e_qq = exp_poly<false>(-(r2 * k.iRqq * k.iRqq));
    const double erf_qq = erf_from_exp_pos(rij * k.iRqq, e_qq, erfcx);
    const double A0 = inv_r3 * (erf_qq - 2.0 * K_INV_SQRT_PI * k.iRqq * rij * e_qq);
    const double sqq = A0 * K_FQQ_AU2 * qi * qj;
    g[0] = sqq * dx; g[1] = sqq * dy; g[2] = sqq * dz;
    e12 += K_FQQ_AU1 * qi * qj * erf_qq * inv_r;
    // charge-dipole, T1 both directions
    const double e_pq = exp_poly<false>(-(r2 * k.iRpq * k.iRpq));
    const double br_pq = erf_from_exp_pos(rij * k.iRpq, e_pq, erfcx) - 2.0 * K_INV_SQRT_PI * k.iRpq * rij * e_pq;
    const double Apq0 = inv_r3 * br_pq;
    const double A1 = inv_r5 * br_pq;
    const double B1 = 4.0 * K_INV_SQRT_PI * k.iRpq3 * inv_r2 * e_pq;
    // w = qj p_i - qi p_j:  g_b += sum_a (A1 (3 r_a r_b - d r^2) - B1 r_a r_b) au2 w_a
    const double w[3] = {qj * pi[0] - qi * pj[0], qj * pi[1] - qi * pj[1], qj * pi[2] - qi * pj[2]};
    const double wr = w[0] * dx + w[1] * dy + w[2] * dz;
#pragma unroll
    for (int b = 0; b < 3; ++b) g[b] += (A1 * (3.0 * wr * rv[b] - w[b] * r2) - B1 * wr * rv[b]) * K_FQQ_AU2;
    const double pjr = pj[0] * dx + pj[1] * dy + pj[2] * dz;
    const double pir = pi[0] * dx + pi[1] * dy + pi[2] * dz;
    e13 += K_FQQ_AU1 * Apq0 * (qj * pir - qi * pjr);
    // dipole-dipole, T2
    const double e_pp = exp_poly<false>(-(r2 * k.iRpp * k.iRpp));
    const double erf_pp = erf_from_exp_pos(rij * k.iRpp, e_pp, erfcx);
    const double App1 = inv_r5 * (erf_pp - 2.0 * K_INV_SQRT_PI * k.iRpp * rij * e_pp);
    const double Bpp1 = 4.0 * K_INV_SQRT_PI * k.iRpp3 * inv_r2 * e_pp;
    const double A2 = inv_r7 * (3.0 * erf_pp - 6.0 * K_INV_SQRT_PI * k.iRpp * rij * e_pp -
                                4.0 * r2 * rij * K_INV_SQRT_PI * k.iRpp3 * e_pp);
    const double B2 = 8.0 * K_INV_SQRT_PI * k.iRpp5 * inv_r2 * e_pp;
    const double pipj = pi[0] * pj[0] + pi[1] * pj[1] + pi[2] * pj[2];
    // sum_ab p_ia p_jb [5 r_a r_b r_c - r^2 (r_a d_bc + r_b d_ca + r_c d_ab)]
    //   = 5 (pi.r)(pj.r) r_c - r^2 ((pi.r) pj_c + (pj.r) pi_c + (pi.pj) r_c)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double t5 = 5.0 * pir * pjr * rv[c] - r2 * (pir * pj[c] + pjr * pi[c] + pipj * rv[c]);
        g[c] -= (A2 * t5 - B2 * pir * pjr * rv[c]) * K_FQQ_AU2;
    }
    e14 -= K_FQQ_AU1 * ((3.0 * pir * pjr - pipj * r2) * App1 - pir * pjr * Bpp1);
}

// atom side of the metal-water terms: non-metal atom a over the metal tail of
// its row, one wavefront; the reaction force of the same pair term (the
// reference's f_atom -= f)
__device__ __forceinline__ void mw_atom_terms(const double* __restrict__ r, const double* __restrict__ charge,
                                              const BoxConst& bc, const int* __restrict__ n_back,
                                              const int* __restrict__ nbr, int stride,
                                              const double* __restrict__ x, int a, int lane,
                                              double* __restrict__ f, const double* __restrict__ erfcx) {
    const int N = bc.N, Nm = bc.Nm;
    double qa = charge[a];
    int nb = n_back[a];
    if (qa == 0.0 || nb == 0) return;
    const int* tail = nbr + (size_t)a * stride + stride - nb;
    double xa = r[a], ya = r[N + a], za = r[2 * N + a];
    double fx = 0.0, fy = 0.0, fz = 0.0, e10 = 0.0, e11 = 0.0;
    for (int k = lane; k < nb; k += 64) {
        int i = tail[k];
        double dx = r[i] - xa, dy = r[N + i] - ya, dz = r[2 * N + i] - za;
        pbc12(bc, dx, dy, dz);
        double r2 = dist2_exact(dx, dy, dz);
        if (!(r2 < bc.rc2)) continue;             // rows reach rc + skin
        double pm[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        double fp[3];
        mw_pair(bc, dx, dy, dz, r2, qa, x[3 * Nm + i], pm, fp, &e10, &e11, erfcx);
        fx += fp[0]; fy += fp[1]; fz += fp[2];
    }
    fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
    if (lane == 0) { f[a] -= fx; f[N + a] -= fy; f[2 * N + a] -= fz; }
}

// The metal atoms' kernels give each metal atom a whole workgroup, its four
// wavefronts splitting the atom's partners: with one wavefront per atom the
// few hundred metal atoms leave most SIMDs with a single wavefront, whose
// transcendental latency nothing hides.  The force on the atom is the
// fixed-order sum of the four wavefront sums.
__device__ __forceinline__ void block_force_add(double fx, double fy, double fz, double (*s_f)[3], int i, int N,
                                                double* __restrict__ f) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
    if (lane == 0) { s_f[wave][0] = fx; s_f[wave][1] = fy; s_f[wave][2] = fz; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += s_f[w][threadIdx.x];
        f[(size_t)threadIdx.x * N + i] += t;
    }
}

// per-block partials: the fixed-order sum over the four wavefronts of each
// of the nval values
template <int NVAL>
__device__ __forceinline__ void block_partials(const double (&vals)[13], double (*s_red)[13],
                                               double* __restrict__ partial, int slot) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NVAL; ++c) {
        double s = wave_sum(vals[c]);
        if (lane == 0) s_red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double s = 0.0;
        if (threadIdx.x < NVAL)
            for (int w = 0; w < 4; ++w) s += s_red[w][threadIdx.x];
        partial[slot * 16 + threadIdx.x] = s;
    }
}

// metal side of the metal-water terms for metal atom i (one workgroup);
// partial[i] = e10, e11, virial
__device__ __forceinline__ void mw_metal_block(const double* __restrict__ r, const double* __restrict__ charge,
                                               const BoxConst& bc, const int* __restrict__ n_front,
                                               const int* __restrict__ nbr, int stride, const double* __restrict__ x,
                                               double* __restrict__ f, double* __restrict__ partial, int i,
                                               const double* s_erfcx, double (*s_red)[13], double (*s_f)[3]) {
    double fx = 0.0, fy = 0.0, fz = 0.0, vals[13];
#pragma unroll
    for (int c = 0; c < 13; ++c) vals[c] = 0.0;
    mw_metal_terms(r, charge, bc, n_front, nbr, stride, x, i, threadIdx.x, fx, fy, fz, vals, s_erfcx);
    block_force_add(fx, fy, fz, s_f, i, bc.N, f);
    block_partials<11>(vals, s_red, partial, i);
}

// atom side of the metal-water terms, workgroup b of nb: the wavefronts
// stride over the non-metal atoms that have metal neighbours (listed by the
// neighbour build; most atoms have none), one atom per wavefront at a time
__device__ __forceinline__ void mw_atom_block(const double* __restrict__ r, const double* __restrict__ charge,
                                              const BoxConst& bc, const int* __restrict__ n_back,
                                              const int* __restrict__ nbr, int stride,
                                              const int* __restrict__ tail_list, const double* __restrict__ x,
                                              double* __restrict__ f, int b, int nb, const double* s_erfcx) {
    const int count = tail_list[0];
    const int nw = nb * 4;
    for (int q = __builtin_amdgcn_readfirstlane(b * 4 + (threadIdx.x >> 6)); q < count; q += nw)
        mw_atom_terms(r, charge, bc, n_back, nbr, stride, x, tail_list[1 + q], threadIdx.x & 63, f, s_erfcx);
}

// workgroups of the atom-side kernel: enough wavefronts to give each listed
// atom its own in one round on this deck size, capped
inline unsigned mw_atom_blocks(int n_nonmetal) {
    const int nb = (n_nonmetal + 3) / 4;
    return (unsigned)(nb < 1 ? 1 : (nb < 2048 ? nb : 2048));
}

// Metal-metal terms on the 64 x 64 tiles of metal blocks (relay_tile.h),
// each unordered pair once.  One workgroup per tile (I, J), I <= J, four
// wavefronts of sixteen diagonals: at diagonal k lane l takes the pair
// i = I*64 + l, j = J*64 + ((l + k) & 63), adds -g to its row sum (registers)
// and +g to its wavefront's column sum for j (LDS; within one step the 64
// lanes meet 64 different j, so no two lanes write one slot).  A diagonal
// tile takes k = 1..31 and half of k = 32: every unordered pair of its block
// once.  The four wavefronts' row and column sums are added in a fixed
// order and stored as the partials P[I][J] (forces on block I from block J)
// and P[J][I] (3 x 64 doubles each; a diagonal tile stores rows + columns);
// k_cpim_mm_sum adds them.  partial (per tile) = e9 (0), e12, e13, e14, virial.
// (162 VGPRs: three workgroups of four wavefronts per compute unit hold all
// 561 tiles of Au2057 at once)
constexpr int kMMWaves = 4;
constexpr int kMMSteps = kTileEdge / kMMWaves;
constexpr int kMMBlk = 3 * kTileEdge;          // doubles of one partial

struct MMShared {
    double j[7][kTileEdge];                 // block J: x, y, z (nm), q, p (3)
    double col[kMMWaves][3][kTileEdge];
    double row[kMMWaves][3][kTileEdge];
    double red[kMMWaves][13];
};

// tile t (the caller has staged the erfcx table and synchronised)
__device__ __forceinline__ void mm_tile_block(const double* __restrict__ r, const BoxConst& bc,
                                              const double* __restrict__ x, int Nb, double* __restrict__ P,
                                              double* __restrict__ partial, int t_slot, const double* s_erfcx,
                                              MMShared& sh) {
    auto& s_j = sh.j;
    auto& s_col = sh.col;
    auto& s_row = sh.row;
    auto& s_red = sh.red;
    const int N = bc.N, Nm = bc.Nm;
    int I, J;
    tile_coords(t_slot, Nb, I, J);
    if (threadIdx.x < kTileEdge) {
        const int j = J * kTileEdge + threadIdx.x;
        const bool ok = j < Nm;
        s_j[0][threadIdx.x] = ok ? r[j] : 0.0;
        s_j[1][threadIdx.x] = ok ? r[N + j] : 0.0;
        s_j[2][threadIdx.x] = ok ? r[2 * N + j] : 0.0;
        s_j[3][threadIdx.x] = ok ? x[3 * Nm + j] : 0.0;
        s_j[4][threadIdx.x] = ok ? x[3 * j] : 0.0;
        s_j[5][threadIdx.x] = ok ? x[3 * j + 1] : 0.0;
        s_j[6][threadIdx.x] = ok ? x[3 * j + 2] : 0.0;
    }
    for (int k = threadIdx.x; k < kMMWaves * 3 * kTileEdge; k += blockDim.x) (&s_col[0][0][0])[k] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = I * kTileEdge + lane;
    const bool iok = i < Nm;
    const double xi = iok ? r[i] : 0.0, yi = iok ? r[N + i] : 0.0, zi = iok ? r[2 * N + i] : 0.0;
    const double qi = iok ? x[3 * Nm + i] : 0.0;
    const double pi[3] = {iok ? x[3 * i] : 0.0, iok ? x[3 * i + 1] : 0.0, iok ? x[3 * i + 2] : 0.0};
    const MMConst mk = mm_const(bc);
    const bool diag = I == J;
    double fr[3] = {0.0, 0.0, 0.0}, e12 = 0.0, e13 = 0.0, e14 = 0.0;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int s = 0; s < kMMSteps; ++s) {
        const int k = wave * kMMSteps + s;
        const int jl = (lane + k) & (kTileEdge - 1);
        bool valid = iok && J * kTileEdge + jl < Nm;
        if (diag) valid = valid && k >= 1 && (k < 32 || (k == 32 && lane < 32));
        if (!valid) continue;
        double dx = s_j[0][jl] - xi, dy = s_j[1][jl] - yi, dz = s_j[2][jl] - zi;
        pbc12(bc, dx, dy, dz);
        dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
        const double pj[3] = {s_j[4][jl], s_j[5][jl], s_j[6][jl]};
        double g[3];
        mm_pair(mk, dx, dy, dz, qi, pi, s_j[3][jl], pj, s_erfcx, g, e12, e13, e14);
        const double rv[3] = {dx, dy, dz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            fr[c] -= g[c];
            s_col[wave][c][jl] += g[c];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) v[a * 3 + b] -= rv[a] * g[b];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) s_row[wave][c][lane] = fr[c];
    double vals[13] = {0.0, e12, e13, e14, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
#pragma unroll
    for (int c = 0; c < 13; ++c) {
        const double t = wave_sum(vals[c]);
        if (lane == 0) s_red[wave][c] = t;
    }
    __syncthreads();
    if (threadIdx.x < kMMBlk) {
        const int c = threadIdx.x >> 6, l = threadIdx.x & 63;
        double rs = 0.0, cs = 0.0;
#pragma unroll
        for (int w = 0; w < kMMWaves; ++w) {
            rs += s_row[w][c][l];
            cs += s_col[w][c][l];
        }
        if (diag) {
            P[((size_t)I * Nb + I) * kMMBlk + threadIdx.x] = rs + cs;
        } else {
            P[((size_t)I * Nb + J) * kMMBlk + threadIdx.x] = rs;
            P[((size_t)J * Nb + I) * kMMBlk + threadIdx.x] = cs;
        }
    } else if (threadIdx.x < kMMBlk + 16) {
        const int c = threadIdx.x - kMMBlk;
        double t = 0.0;
        if (c < 13)
            for (int w = 0; w < kMMWaves; ++w) t += s_red[w][c];
        partial[t_slot * 16 + c] = t;
    }
}

// The CPIM force terms in one launch: workgroups [0, n_tiles) the
// metal-metal tiles (partials for k_cpim_mm_sum, energies to part_mm),
// [n_tiles, n_tiles + n_metal) the metal side of the metal-water terms (one
// per metal atom; forces added to f, energies to part_mw), then n_atom
// workgroups for the atom side (forces added to the non-metal atoms).  The
// three write different things (the tiles touch no force), so they share the
// compute units instead of running one after the other.
__global__ void __launch_bounds__(kMMWaves * 64)
k_cpim_terms(const double* __restrict__ r, const double* __restrict__ charge, BoxConst bc,
             const double* __restrict__ x, int Nb, double* __restrict__ P, double* __restrict__ part_mm,
             int n_tiles, const int* __restrict__ n_front, const int* __restrict__ n_back,
             const int* __restrict__ nbr, int stride, const int* __restrict__ tail_list, double* __restrict__ f,
             double* __restrict__ part_mw, int n_metal, int n_atom, int n_sc, const double* __restrict__ isd,
             double* __restrict__ f_sc, double* __restrict__ part_sc) {
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    __shared__ MMShared sh;
    int b = blockIdx.x;
    if (b >= n_tiles + n_metal + n_atom) {
        // the Sutton-Chen forces of four metal atoms, written to f_sc
        sc_force_block(b - n_tiles - n_metal - n_atom, r, bc, n_back, nbr, stride, isd, f_sc, bc.Nm, 0, part_sc,
                       reinterpret_cast<double (*)[10]>(&sh.red[0][0]));
        return;
    }
    load_erfcx_table(s_erfcx);
    __syncthreads();
    if (b < n_tiles) {
        mm_tile_block(r, bc, x, Nb, P, part_mm, b, s_erfcx, sh);
        return;
    }
    b -= n_tiles;
    if (b < n_metal) {
        mw_metal_block(r, charge, bc, n_front, nbr, stride, x, f, part_mw, b, s_erfcx, sh.red,
                       reinterpret_cast<double (*)[3]>(&sh.row[0][0][0]));
        return;
    }
    b -= n_metal;
    mw_atom_block(r, charge, bc, n_back, nbr, stride, tail_list, x, f, b, n_atom, s_erfcx);
}

// The metal atoms' CPIM forces from the tiles' partials: one workgroup of 256
// threads per block B of 64 metal atoms; thread (c, l), c < 3, sums
// P[B][S] over S = 0 .. Nb-1 in order for atom B*64 + l, adds the external
// field on the induced charge (cpff.c:930-948) and extra (the Sutton-Chen
// force, 3 x N_m, or NULL), and adds the total to f.  The self terms
// (cpff.c:727-742) and the field energies of the block's charges and dipoles
// go to partial[slot] (e9, e12, e13 = 0, e14, virial 0).
__global__ void __launch_bounds__(256)
k_cpim_mm_sum(const double* __restrict__ r, BoxConst bc, const double* __restrict__ x, int Nb,
              const double* __restrict__ P, const double* __restrict__ extra, double* __restrict__ f,
              double* __restrict__ partial, int slot0, double* __restrict__ hist, int n_mat) {
    const int N = bc.N, Nm = bc.Nm, B = blockIdx.x;
    // (hist non-NULL: the solution's copy for the warm-start predictor's
    // history, in the same launch)
    if (hist)
        for (int k = B * blockDim.x + threadIdx.x; k < n_mat; k += gridDim.x * blockDim.x) hist[k] = x[k];
    if (threadIdx.x < kMMBlk) {
        const int c = threadIdx.x >> 6, l = threadIdx.x & 63, i = B * kTileEdge + l;
        // (the loads of a chunk issued together, the sum in order)
        double t = 0.0;
        for (int S0 = 0; S0 < Nb; S0 += 16) {
            double pv[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
                pv[q] = S0 + q < Nb ? P[((size_t)B * Nb + S0 + q) * kMMBlk + threadIdx.x] : 0.0;
#pragma unroll
            for (int q = 0; q < 16; ++q) t += pv[q];
        }
        if (i < Nm) {
            t += bc.efield[c] * x[3 * Nm + i];
            if (extra) t += extra[(size_t)c * Nm + i];
            f[(size_t)c * N + i] += t;
        }
    }
    // self and field energies of the block's atoms (the first wavefront)
    if (threadIdx.x < kTileEdge) {
        double e9 = 0.0, e12 = 0.0, e14 = 0.0;
        const int i = B * kTileEdge + (int)threadIdx.x;
        if (i < Nm) {
            const double qi = x[3 * Nm + i];
            const double pi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
            // (inv_capac, inv_polar in atomic units, FQQ in MD units)
            e12 = 0.5 * K_FQQ * qi * qi * (bc.inv_capac * K_NM2BOHR);
            e14 = 0.5 * K_FQQ * (pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2]) * (bc.inv_polar * K_NM2BOHR);
            e9 = -(bc.efield[0] * qi * r[i] + bc.efield[1] * qi * r[N + i] + bc.efield[2] * qi * r[2 * N + i]);
            e9 -= bc.efield[0] * pi[0] + bc.efield[1] * pi[1] + bc.efield[2] * pi[2];
        }
        e9 = wave_sum(e9);
        e12 = wave_sum(e12);
        e14 = wave_sum(e14);
        if (threadIdx.x < 16) {
            const int c = threadIdx.x;
            partial[(size_t)(slot0 + B) * 16 + c] = c == 0 ? e9 : c == 1 ? e12 : c == 3 ? e14 : 0.0;
        }
    }
}

}  // namespace

// b (n_mat) from the current positions
int cpim_rhs(capac_ctx* c) {
    int Nm = c->bc.Nm;
    if (Nm == 0) return 0;
    if (c->bc.n_np > Nm) {
        capac_set_error("more nanoparticles (%d) than metal atoms (%d)", c->bc.n_np, Nm);
        return CAPAC_ERR_ARG;
    }
    k_rhs<<<Nm, 256, 0, c->stream>>>(c->r, c->charge, c->bc, c->n_front, c->nbr,
                                               c->nb_stride, c->np_charge, c->b);
    CHECK_HIP(hipGetLastError());
    return 0;
}

static int assemble_finish(capac_ctx* c);

// b and the relay of a force call: one launch on the FP32 path's TILES
// layout, else cpim_rhs and cpim_assemble
int cpim_rhs_assemble(capac_ctx* c) {
    const int Nm = c->bc.Nm;
    if (Nm == 0) return 0;
    if (c->precision != CAPAC_PREC_FP32_REFINED || c->relay_layout != CAPAC_LAYOUT_TILES) {
        int rc = cpim_rhs(c);
        return rc ? rc : cpim_assemble(c);
    }
    if (c->bc.n_np > Nm) {
        capac_set_error("more nanoparticles (%d) than metal atoms (%d)", c->bc.n_np, Nm);
        return CAPAC_ERR_ARG;
    }
    const int n_tiles = (int)tile_count(c->relay_nb);
    const int n_sc = (Nm + 3) / 4;          // the Sutton-Chen densities ride along (sc_in_force_launches)
    k_matrix_terms<<<1 + n_tiles + Nm + n_sc, 256, 0, c->stream>>>(
        c->r, c->charge, c->bc, c->n_front, c->nbr, c->nb_stride, c->np_charge, c->b, c->relay_nb, n_tiles,
        reinterpret_cast<float4*>(c->relay32), c->metal_u, c->n_back, c->isd);
    CHECK_HIP(hipGetLastError());
    return assemble_finish(c);
}

int cpim_assemble(capac_ctx* c) {
    int Nm = c->bc.Nm;
    if (Nm == 0) return 0;
    // the FP64 path stores nothing: its SpMV forms each block from the positions
    dim3 grid((Nm + 255) / 256, Nm);
    if (c->precision == CAPAC_PREC_FP32_REFINED) {
        if (c->relay_layout == CAPAC_LAYOUT_TILES)
            k_assemble32_tiles<<<(unsigned)tile_count(c->relay_nb), 256, 0, c->stream>>>(
                c->r, c->bc, c->relay_nb, reinterpret_cast<float4*>(c->relay32));
        else
            k_assemble32<<<grid, 256, 0, c->stream>>>(c->r, c->bc, c->relay_ld, c->relay32);
        k_metal_unwrap<<<1, 1024, 0, c->stream>>>(c->r, c->bc, c->metal_u);
    }
    return assemble_finish(c);
}

// the diagonal and the block preconditioner after the relay's launch
static int assemble_finish(capac_ctx* c) {
    // the Jacobi diagonal holds constants only (1/alpha, -1/c, 1): formed once
    // and again only when those constants change (fitting parameters)
    if (!c->diag_set || c->diag_polar != c->bc.inv_polar || c->diag_capac != c->bc.inv_capac) {
        k_diag<<<(c->n_mat + 255) / 256, 256, 0, c->stream>>>(c->bc, c->n_mat, c->diag);
        c->diag_set = 1;
        c->diag_polar = c->bc.inv_polar;
        c->diag_capac = c->bc.inv_capac;
        c->bj_valid = 0;                // the blocks' diagonal changed with them
    }
    // block-Jacobi inverses: every bj_every assemblies (the blocks change
    // little from one MD step to the next, and the preconditioner only steers
    // the iterations: the stop test is the same)
    if (c->precision == CAPAC_PREC_FP32_REFINED && c->bj_inv && c->bj_every > 0 &&
        (!c->bj_valid || ++c->bj_age >= c->bj_every)) {
        if (int rc = bj_build(c)) return rc;
        c->bj_valid = 1;
        c->bj_age = 0;
    }
    CHECK_HIP(hipGetLastError());
    return 0;
}

int cpim_dense_expand(capac_ctx* c, double* dense) {
    long n4 = 4L * c->bc.Nm;
    unsigned blocks = (unsigned)((n4 * n4 + 255) / 256);
    k_dense<<<blocks, 256, 0, c->stream>>>(c->precision == CAPAC_PREC_FP32_REFINED ? c->relay32 : NULL,
                                           c->relay_layout == CAPAC_LAYOUT_TILES, c->relay_nb, c->r, c->bc,
                                           c->relay_ld, dense);
    CHECK_HIP(hipGetLastError());
    return 0;
}

int cpim_apply64(capac_ctx* c, const double* x, const double* b, double* out, double* part,
                 const SolveScalars* guard, int* n_part) {
    const int Nm = c->bc.Nm, n_np = c->bc.n_np;
    if (n_part) *n_part = 0;
    if (Nm == 0) return 0;
    if (c->relay_layout == CAPAC_LAYOUT_TILES && c->relay_part) {
        // (the FP32 product's partials are free here: no product is in flight)
        const int Nb = tile_blocks(Nm);
        k_apply64_tiles<<<(unsigned)tile_count(Nb), 256, 0, c->stream>>>(c->r, c->bc, x, Nb, c->relay_part, guard);
        k_apply64_sum<<<Nb + n_np, 256, 0, c->stream>>>(c->bc, c->np_of_metal, c->np_first, c->np_last, x, b, Nb,
                                                        c->relay_part, out, part, guard);
        if (n_part) *n_part = Nb + n_np;
    } else {
        k_apply64<<<Nm + n_np, 256, 0, c->stream>>>(c->r, c->bc, c->np_of_metal, c->np_first, c->np_last, x, b,
                                                    out, part, guard);
        if (n_part) *n_part = Nm + n_np;
    }
    CHECK_HIP(hipGetLastError());
    return 0;
}

// the metal-metal terms for x: the tile launch and the per-block sums
// (adding extra, 3 x N_m or NULL, with them); reduce: queue the energies'
// reduction (red[48..61))
// mw: the metal-water terms in the same launch; sc: the Sutton-Chen forces
// too (into extra, which the block sums then add)
static int cpim_mm_launch(capac_ctx* c, const double* x, double* extra, bool reduce, bool mw,
                          bool sc = false) {
    const int Nm = c->bc.Nm, N = c->bc.N;
    const int Nb = tile_blocks(Nm);
    const int n_tiles = (int)tile_count(Nb);
    const int n_metal = mw ? Nm : 0;
    const int n_atom = mw && N > Nm ? (int)mw_atom_blocks(N - Nm) : 0;
    const int n_sc = sc && extra ? (Nm + 3) / 4 : 0;
    double* part = partial_seg(c, SEG_MM);
    k_cpim_terms<<<n_tiles + n_metal + n_atom + n_sc, kMMWaves * 64, 0, c->stream>>>(
        c->r, c->charge, c->bc, x, Nb, c->mm_part, part, n_tiles, c->n_front, c->n_back, c->nbr, c->nb_stride,
        c->tail_list, c->f, partial_seg(c, SEG_MW), n_metal, n_atom, n_sc, c->isd, extra,
        partial_seg(c, SEG_SC));
    // a force call of an MD step also copies x into the predictor's history
    double* hist = reduce && mw && x == c->x ? c->hist_dst : NULL;
    if (hist) {
        c->hist_copied = hist;          // (asked once per step)
        c->hist_dst = NULL;
    }
    k_cpim_mm_sum<<<Nb, 256, 0, c->stream>>>(c->r, c->bc, x, Nb, c->mm_part, extra, c->f, part, n_tiles, hist,
                                             c->n_mat);
    CHECK_HIP(hipGetLastError());
    if (!reduce) return 0;
    if (mw) {
        int rc = reduce_energies(c, SEG_MW, Nm, 11, 32);
        if (rc) return rc;
    }
    if (n_sc) {
        int rc = reduce_energies(c, SEG_SC, n_sc, 10, 16);
        if (rc) return rc;
    }
    return reduce_energies(c, SEG_MM, n_tiles + Nb, 13, 48);
}

// adds the CPIM forces for x (device) to c->f;
// red[32..43) = pot10, pot11, vir; red[48..61) = pot9, pot12, pot13, pot14, vir
int cpim_forces_x(capac_ctx* c, const double* x, double* f_sc) {
    if (c->bc.Nm == 0) return 0;
    return cpim_mm_launch(c, x, f_sc, true, true, sc_in_force_launches(c));
}

// the force call's Sutton-Chen passes ride in its RHS/relay launch
// (densities) and CPIM force launch (forces): the FP32 path's TILES layout
bool sc_in_force_launches(const capac_ctx* c) {
    return c->bc.Nm > 0 && c->n_mat > 0 && c->precision == CAPAC_PREC_FP32_REFINED &&
           c->relay_layout == CAPAC_LAYOUT_TILES;
}

// (in a force call the Sutton-Chen force waits in f_sc)
int cpim_forces(capac_ctx* c) { return cpim_forces_x(c, c->x, c->f_sc); }

// the metal-metal CPIM force launches alone, adding to c->f (for kernel timing)
int cpim_mm_only(capac_ctx* c) {
    if (c->bc.Nm == 0) return 0;
    return cpim_mm_launch(c, c->x, NULL, false, false);
}
