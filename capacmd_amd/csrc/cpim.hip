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

namespace {

// ------------------------------------------------------------------ RHS
// One workgroup per metal atom over the non-metal front of its row (four
// wavefronts split the row, as in the metal-side force kernels).
__global__ void __launch_bounds__(256)
k_rhs(const double* __restrict__ r, const double* __restrict__ charge, BoxConst bc,
      const int* __restrict__ n_front, const int* __restrict__ nbr, int stride,
      const double* __restrict__ np_charge, double* __restrict__ b) {
    main_stream_priority();
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    __shared__ double s_w[4][4];
    const int N = bc.N, Nm = bc.Nm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x;
    load_erfcx_table(s_erfcx);
    __syncthreads();
    if (i >= Nm) return;
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
__global__ void __launch_bounds__(256)
k_assemble32_tiles(const double* __restrict__ r, BoxConst bc, int Nb, float4* __restrict__ tiles) {
    main_stream_priority();
    const int Nm = bc.Nm;
    int I, J;
    tile_coords(blockIdx.x, Nb, I, J);
    float4* dst = tiles + (size_t)blockIdx.x * kTileSlots;
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

// Metal separations from metal atom 0, d_j = pbc_12(r_j - r_0) in bohr, and
// whether they are compact: when every component's spread max_j d_j -
// min_j d_j is below half the box edge, every pair's minimum-image
// separation is d_j - d_i (pbc_12 leaves a difference inside half the box
// alone), and the FP32 product applies its blocks from d without the
// per-pair image test.  One workgroup; the flag lands at u[4 Nm].
__global__ void __launch_bounds__(1024)
k_metal_unwrap(const double* __restrict__ r, BoxConst bc, double* __restrict__ u) {
    main_stream_priority();
    __shared__ double s_lo[3][16], s_hi[3][16];
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

// Metal-metal terms, all pairs, force on i only (cpff.c:762-927), plus the
// self terms (cpff.c:727-742) and the external field on the induced charge
// and dipole (cpff.c:930-948); j split over 256 threads (t = 0..255).  Each
// erf comes from the Gaussian the same term needs anyway (erf_from_exp, with
// the erfcx table staged in LDS), not from a library erf of its own.
// vals = e9, e12, e13, e14, virial
__device__ __forceinline__ void mm_terms(const double* __restrict__ r, const BoxConst& bc,
                                         const double* __restrict__ x, const double* __restrict__ erfcx,
                                         int i, int t,
                                         double& fx, double& fy, double& fz, double vals[13]) {
    const int N = bc.N, Nm = bc.Nm;
    double e9 = 0.0, e12 = 0.0, e13 = 0.0, e14 = 0.0, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
    const double qi = x[3 * Nm + i];
    const double pi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const double iRqq = bc.inv_R_qq, iRpq = bc.inv_R_pq, iRpp = bc.inv_R_pp;
    const double iRpq3 = iRpq * iRpq * iRpq, iRpp3 = iRpp * iRpp * iRpp;
    const double iRpp5 = iRpp3 * iRpp * iRpp;
    for (int j = t; j < Nm; j += 256) {
        if (j == i) continue;
        double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
        pbc12(bc, dx, dy, dz);
        dx *= K_NM2BOHR; dy *= K_NM2BOHR; dz *= K_NM2BOHR;
        double r2 = dist2_exact(dx, dy, dz);
        double inv_r = rsqrt_pos(r2), inv_r2 = inv_r * inv_r;
        double rij = r2 * inv_r;
        double inv_r3 = inv_r2 * inv_r, inv_r5 = inv_r3 * inv_r2, inv_r7 = inv_r5 * inv_r2;
        double qj = x[3 * Nm + j];
        double pj[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
        double rv[3] = {dx, dy, dz};
        double g[3] = {0.0, 0.0, 0.0};
        // charge-charge, T0
        double e_qq = exp_poly<false>(-(r2 * iRqq * iRqq));
        double erf_qq = erf_from_exp(rij * iRqq, e_qq, erfcx);
        double A0 = inv_r3 * (erf_qq - 2.0 * K_INV_SQRT_PI * iRqq * rij * e_qq);
        double sqq = A0 * K_FQQ_AU2 * qi * qj;
        g[0] += sqq * dx; g[1] += sqq * dy; g[2] += sqq * dz;
        e12 += 0.5 * K_FQQ_AU1 * qi * qj * erf_qq * inv_r;
        // charge-dipole, T1 both directions
        double e_pq = exp_poly<false>(-(r2 * iRpq * iRpq));
        double br_pq = erf_from_exp(rij * iRpq, e_pq, erfcx) - 2.0 * K_INV_SQRT_PI * iRpq * rij * e_pq;
        double Apq0 = inv_r3 * br_pq;
        double A1 = inv_r5 * br_pq;
        double B1 = 4.0 * K_INV_SQRT_PI * iRpq3 * inv_r2 * e_pq;
        // w = qj p_i - qi p_j:  f_b += sum_a (A1 (3 r_a r_b - d r^2) - B1 r_a r_b) au2 w_a
        double w[3] = {qj * pi[0] - qi * pj[0], qj * pi[1] - qi * pj[1], qj * pi[2] - qi * pj[2]};
        double wr = w[0] * dx + w[1] * dy + w[2] * dz;
#pragma unroll
        for (int b = 0; b < 3; ++b)
            g[b] += (A1 * (3.0 * wr * rv[b] - w[b] * r2) - B1 * wr * rv[b]) * K_FQQ_AU2;
        double pjr = pj[0] * dx + pj[1] * dy + pj[2] * dz;
        double pir = pi[0] * dx + pi[1] * dy + pi[2] * dz;
        e13 += 0.5 * K_FQQ_AU1 * Apq0 * (qj * pir - qi * pjr);
        // dipole-dipole, T2
        double e_pp = exp_poly<false>(-(r2 * iRpp * iRpp));
        double erf_pp = erf_from_exp(rij * iRpp, e_pp, erfcx);
        double App1 = inv_r5 * (erf_pp - 2.0 * K_INV_SQRT_PI * iRpp * rij * e_pp);
        double Bpp1 = 4.0 * K_INV_SQRT_PI * iRpp3 * inv_r2 * e_pp;
        double A2 = inv_r7 * (3.0 * erf_pp - 6.0 * K_INV_SQRT_PI * iRpp * rij * e_pp -
                              4.0 * r2 * rij * K_INV_SQRT_PI * iRpp3 * e_pp);
        double B2 = 8.0 * K_INV_SQRT_PI * iRpp5 * inv_r2 * e_pp;
        double pipj = pi[0] * pj[0] + pi[1] * pj[1] + pi[2] * pj[2];
        // sum_ab p_ia p_jb [5 r_a r_b r_c - r^2 (r_a d_bc + r_b d_ca + r_c d_ab)]
        //   = 5 (pi.r)(pj.r) r_c - r^2 ((pi.r) pj_c + (pj.r) pi_c + (pi.pj) r_c)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double t5 = 5.0 * pir * pjr * rv[c] - r2 * (pir * pj[c] + pjr * pi[c] + pipj * rv[c]);
            g[c] -= (A2 * t5 - B2 * pir * pjr * rv[c]) * K_FQQ_AU2;
        }
        e14 -= 0.5 * K_FQQ_AU1 * ((3.0 * pir * pjr - pipj * r2) * App1 - pir * pjr * Bpp1);
        fx -= g[0]; fy -= g[1]; fz -= g[2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) v[a * 3 + b] -= 0.5 * rv[a] * g[b];
    }
    if (t == 0) {
        // self terms (inv_capac, inv_polar in atomic units, FQQ in MD units)
        e12 += 0.5 * K_FQQ * qi * qi * (bc.inv_capac * K_NM2BOHR);
        e14 += 0.5 * K_FQQ * (pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2]) * (bc.inv_polar * K_NM2BOHR);
        // external field on the induced charge (force) and charge + dipole (energy)
        fx += bc.efield[0] * qi; fy += bc.efield[1] * qi; fz += bc.efield[2] * qi;
        e9 -= bc.efield[0] * qi * xi + bc.efield[1] * qi * yi + bc.efield[2] * qi * zi;
        e9 -= bc.efield[0] * pi[0] + bc.efield[1] * pi[1] + bc.efield[2] * pi[2];
    }
    vals[0] = e9; vals[1] = e12; vals[2] = e13; vals[3] = e14;
#pragma unroll
    for (int c = 0; c < 9; ++c) vals[4 + c] = v[c];
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
// extra (3 x Nm, or NULL): a force the same write adds (the Sutton-Chen force)
__device__ __forceinline__ void block_force_add(double fx, double fy, double fz, double (*s_f)[3], int i, int N,
                                                double* __restrict__ f, const double* __restrict__ extra = NULL,
                                                int Nm = 0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
    if (lane == 0) { s_f[wave][0] = fx; s_f[wave][1] = fy; s_f[wave][2] = fz; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += s_f[w][threadIdx.x];
        if (extra) t += extra[(size_t)threadIdx.x * Nm + i];
        f[(size_t)threadIdx.x * N + i] += t;
    }
}

// per-block partials: the fixed-order sum over the four wavefronts of each
// of the nval values
template <int NVAL>
__device__ __forceinline__ void block_partials(const double (&vals)[13], double (*s_red)[13],
                                               double* __restrict__ partial) {
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
        partial[blockIdx.x * 16 + threadIdx.x] = s;
    }
}

// metal side of the metal-water terms: one workgroup per metal atom;
// partial = e10, e11, virial
__global__ void __launch_bounds__(256)
k_cpim_mw_metal(const double* __restrict__ r, const double* __restrict__ charge, BoxConst bc,
                const int* __restrict__ n_front, const int* __restrict__ nbr, int stride,
                const double* __restrict__ x, double* __restrict__ f, double* __restrict__ partial) {
    __shared__ double s_red[4][13];
    __shared__ double s_f[4][3];
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    load_erfcx_table(s_erfcx);
    __syncthreads();
    const int i = blockIdx.x;
    double fx = 0.0, fy = 0.0, fz = 0.0, vals[13];
#pragma unroll
    for (int c = 0; c < 13; ++c) vals[c] = 0.0;
    mw_metal_terms(r, charge, bc, n_front, nbr, stride, x, i, threadIdx.x, fx, fy, fz, vals, s_erfcx);
    block_force_add(fx, fy, fz, s_f, i, bc.N, f);
    block_partials<11>(vals, s_red, partial);
}

// atom side of the metal-water terms: the wavefronts stride over the
// non-metal atoms that have metal neighbours (listed by the neighbour build;
// most atoms have none), one atom per wavefront at a time
__global__ void __launch_bounds__(256)
k_cpim_mw_atom(const double* __restrict__ r, const double* __restrict__ charge, BoxConst bc,
               const int* __restrict__ n_back, const int* __restrict__ nbr, int stride,
               const int* __restrict__ tail_list, const double* __restrict__ x, double* __restrict__ f) {
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    load_erfcx_table(s_erfcx);
    __syncthreads();
    const int count = tail_list[0];
    const int nw = gridDim.x * 4;
    for (int q = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6)); q < count; q += nw)
        mw_atom_terms(r, charge, bc, n_back, nbr, stride, x, tail_list[1 + q], threadIdx.x & 63, f, s_erfcx);
}

// workgroups of the atom-side kernel: enough wavefronts to give each listed
// atom its own in one round on this deck size, capped
inline unsigned mw_atom_blocks(int n_nonmetal) {
    const int nb = (n_nonmetal + 3) / 4;
    return (unsigned)(nb < 1 ? 1 : (nb < 2048 ? nb : 2048));
}

// metal-metal terms: one workgroup per metal atom;
// partial = e9, e12, e13, e14, virial
__global__ void __launch_bounds__(256)
k_cpim_mm(const double* __restrict__ r, BoxConst bc, const double* __restrict__ x,
          double* __restrict__ f, double* __restrict__ partial, const double* __restrict__ f_sc) {
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    __shared__ double s_red[4][13];
    __shared__ double s_f[4][3];
    load_erfcx_table(s_erfcx);
    __syncthreads();
    const int i = blockIdx.x;
    double fx = 0.0, fy = 0.0, fz = 0.0, vals[13];
#pragma unroll
    for (int c = 0; c < 13; ++c) vals[c] = 0.0;
    mm_terms(r, bc, x, s_erfcx, i, threadIdx.x, fx, fy, fz, vals);
    block_force_add(fx, fy, fz, s_f, i, bc.N, f, f_sc, bc.Nm);
    block_partials<13>(vals, s_red, partial);
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
                 const SolveScalars* guard) {
    const int groups = c->bc.Nm + c->bc.n_np;
    if (c->bc.Nm == 0) return 0;
    k_apply64<<<groups, 256, 0, c->stream>>>(c->r, c->bc, c->np_of_metal, c->np_first, c->np_last, x, b,
                                             out, part, guard);
    CHECK_HIP(hipGetLastError());
    return 0;
}

// adds the CPIM forces for x (device) to c->f;
// red[32..43) = pot10, pot11, vir; red[48..61) = pot9, pot12, pot13, pot14, vir
int cpim_forces_x(capac_ctx* c, const double* x, const double* f_sc) {
    int Nm = c->bc.Nm, N = c->bc.N;
    if (Nm == 0) return 0;
    int nblk = Nm;                      // one workgroup per metal atom
    k_cpim_mw_metal<<<nblk, 256, 0, c->stream>>>(c->r, c->charge, c->bc, c->n_front, c->nbr,
                                                 c->nb_stride, x, c->f, partial_seg(c, SEG_MW));
    CHECK_HIP(hipGetLastError());
    int rc = reduce_partials(c, SEG_MW, nblk, 11, 32);
    if (rc) return rc;
    if (N > Nm)
        k_cpim_mw_atom<<<mw_atom_blocks(N - Nm), 256, 0, c->stream>>>(c->r, c->charge, c->bc, c->n_back,
                                                                        c->nbr, c->nb_stride, c->tail_list, x,
                                                                        c->f);
    k_cpim_mm<<<nblk, 256, 0, c->stream>>>(c->r, c->bc, x, c->f, partial_seg(c, SEG_MM), f_sc);
    CHECK_HIP(hipGetLastError());
    return reduce_partials(c, SEG_MM, nblk, 13, 48);
}

// (in a force call the Sutton-Chen force waits in f_sc)
int cpim_forces(capac_ctx* c) { return cpim_forces_x(c, c->x, c->f_sc); }

// the metal-metal CPIM force kernel alone, adding to c->f (for kernel timing)
int cpim_mm_only(capac_ctx* c) {
    int Nm = c->bc.Nm;
    if (Nm == 0) return 0;
    k_cpim_mm<<<Nm, 256, 0, c->stream>>>(c->r, c->bc, c->x, c->f, partial_seg(c, SEG_MM), NULL);
    CHECK_HIP(hipGetLastError());
    return 0;
}
