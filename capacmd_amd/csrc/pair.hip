// Non-bonded pair forces: shifted-force LJ + damped-shifted-force Coulomb
// (funct 1), Buckingham (2), Morse (3) and erf-damped vdW (4), the
// reference's mpi_nonb_pair + compute_nonbonded + compute_buckingham +
// compute_morse + compute_erf_vdw (force.c:1402-1540, 710-816, 639-703,
// 591-632).
//
// One wavefront per atom i over its full neighbour row (front and back
// segments); lanes take consecutive neighbours, so the index stream is
// coalesced and the j-position gathers hit L2.  The force on i is a fixed-order
// wavefront reduction written once (no atomics); each pair is evaluated from
// both ends, so pair energies and the virial carry a factor 1/2.
// FP64-VALU bound (erfc + exp per charged pair).
#include "capac_internal.h"
#include "erfc_fast.h"

namespace {

constexpr int kWavesPerBlock = 4;

// what a lane accumulates: the force on the current atom i, and (over all
// of the lane's pairs) the two energies and the virial, symmetric for central
// pair forces, so six components: xx, xy, xz, yy, yz, zz
struct PairOut {
    double fx, fy, fz, e7, e8, e9, v[6];
};

// The reference's expressions with each division by r, r^2, r^6 or the erf
// width replaced by a product with 1/r from one reciprocal square root per
// pair, and r = r^2 (1/r); the shifted-force LJ constants of the pair type
// come from the table (results move by rounding only).
__device__ __forceinline__ void pair_term(const BoxConst& bc, const PairParam& pp, double qi, double qj,
                                          double dx, double dy, double dz, double r2, PairOut& o,
                                          const double* __restrict__ erfcx_coef) {
    double flj = 0.0, fqq = 0.0;
    const double inv_r = rsqrt(r2);
    const double rij = r2 * inv_r;
    const double inv_r2 = inv_r * inv_r;
    if (pp.funct == 1) {
        if (pp.c6 != 0.0 || pp.c12 != 0.0) {
            const double inv_r6 = inv_r2 * inv_r2 * inv_r2;
            const double cr12 = pp.c12 * (inv_r6 * inv_r6), cr6 = pp.c6 * inv_r6;
            if (bc.use_vdw == 1) {
                // shifted force LJ (force.c:748-757)
                double f_r = 12.0 * cr12 - 6.0 * cr6;
                flj = (f_r - pp.lj_fc) * inv_r2;
                o.e8 += (cr12 - cr6) + (rij - bc.rc) * pp.lj_fc_rc - pp.lj_ec;
            } else {
                flj = (2.0 * cr12 - cr6) * 6.0 * inv_r2;
                o.e8 += cr12 - cr6;
            }
        }
        if (qi != 0.0 && qj != 0.0) {
            double qq = K_FQQ * qi * qj;
            if (bc.use_coulomb == 1) {
                // damped shifted force (force.c:787-795)
                double ex = exp(-bc.alpha * bc.alpha * r2);
                double erfc_a = erfc_from_exp(bc.alpha * rij, ex, erfcx_coef);
                fqq = qq * ((erfc_a * inv_r2 + bc.a2_sqrtPI * ex * inv_r) - bc.wolfConst) * inv_r;
                o.e7 += qq * (erfc_a * inv_r - bc.erfc_arCut + bc.wolfConst * (rij - bc.rc));
            } else {
                fqq = qq * inv_r2 * inv_r;
                o.e7 += qq * inv_r;
            }
        }
    } else if (pp.funct == 4) {
        if (pp.c6 != 0.0 || pp.c12 != 0.0) {
            // erf-damped vdW (force.c:602-628)
            const double inv_r6 = inv_r2 * inv_r2 * inv_r2;
            const double cr12 = pp.c12 * (inv_r6 * inv_r6), cr6 = pp.c6 * inv_r6;
            double rR = rij * pp.inv_width;
            double erf_rR = erf(rR);
            o.e8 += (cr12 - cr6) * erf_rR;
            flj = ((12.0 * cr12 - 6.0 * cr6) * inv_r * erf_rR -
                   (cr12 - cr6) * 2.0 * K_INV_SQRT_PI * exp(-rR * rR) * pp.inv_width) * inv_r;
        }
    } else if (pp.funct == 2) {
        // Buckingham (force.c:674-703): A e^{-B r} - C6 / r^6, no Coulomb term
        const double inv_r6 = inv_r2 * inv_r2 * inv_r2;
        const double e = exp(-pp.b * rij);
        flj = pp.a * e * pp.b * inv_r - 6.0 * pp.c6 * (inv_r6 * inv_r2);
        o.e8 += pp.a * e - pp.c6 * inv_r6;
    } else if (pp.funct == 3) {
        // Morse (force.c:639-668): D (e^{-2a(r-R)} - 2 e^{-a(r-R)}), D in a,
        // the width a in b, R in c6; no Coulomb term
        const double e = exp(-pp.b * (rij - pp.c6));
        const double e2 = e * e;
        flj = 2.0 * pp.a * pp.b * (e2 - e) * inv_r;
        o.e8 += pp.a * (e2 - 2.0 * e);
    }
    double fs = flj + fqq;               // f_j = fs * r_ij, f_i = -f_j
    double fjx = fs * dx, fjy = fs * dy, fjz = fs * dz;
    o.fx -= fjx; o.fy -= fjy; o.fz -= fjz;
    o.v[0] -= dx * fjx; o.v[1] -= dx * fjy; o.v[2] -= dx * fjz;
    o.v[3] -= dy * fjy; o.v[4] -= dy * fjz; o.v[5] -= dz * fjz;
}

// Registers decide this kernel's speed: at 2 wavefronts per SIMD (207
// VGPRs) the transcendental chains were exposed; per-lane energy and virial
// totals (no per-atom copies), a symmetric 6-component virial, the atom's
// data in scalar registers and a cap of 3 wavefronts per SIMD (two VGPRs
// spill) took the full-grid launch from 241 to 174 us on Au561 + water.
__global__ void __launch_bounds__(256, 3)
k_pair(const double* __restrict__ r, const double* __restrict__ charge, const int* __restrict__ type,
       const PairParam* __restrict__ table, BoxConst bc, const int* __restrict__ n_front,
       const int* __restrict__ n_back, const int* __restrict__ nbr, int stride,
       double* __restrict__ f, int accumulate, int field, double* __restrict__ partial) {
    __shared__ PairParam s_tab[64];
    __shared__ double s_red[kWavesPerBlock][12];
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    const int N = bc.N, nt = bc.n_types;
    for (int k = threadIdx.x; k < nt * nt && k < 64; k += blockDim.x) s_tab[k] = table[k];
    load_erfcx_table(s_erfcx);
    __syncthreads();
    const int lane = threadIdx.x & (CAPAC_WAVE - 1);
    // the atom index is the same on every lane: keep it (and what is loaded
    // through it) in scalar registers
    const int first = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + threadIdx.x / CAPAC_WAVE);
    PairOut o = {};
    for (int i = first; i < N; i += gridDim.x * kWavesPerBlock) {
        double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
        double qi = charge[i];
        int ti = type[i] * nt;
        int nf = n_front[i], nb = n_back[i];
        const int* row = nbr + (size_t)i * stride;
        o.fx = 0.0; o.fy = 0.0; o.fz = 0.0;
        const int total = nf + nb;
        // two-stage software pipeline over the lane's pairs: the neighbour
        // index two rounds ahead and the neighbour's data one round ahead are
        // in flight while the current pair is evaluated, so a wavefront does
        // not wait out two dependent loads every round (past the row's end
        // the index is i itself, a valid address whose data go unused)
        auto nbr_at = [&](int kk) { return kk < total ? (kk < nf ? row[kk] : row[stride - nb + (kk - nf)]) : i; };
        int k = lane;
        const int j0 = nbr_at(k);
        int j1 = nbr_at(k + CAPAC_WAVE);
        double x0 = r[j0], y0 = r[N + j0], z0 = r[2 * N + j0], q0 = charge[j0];
        int t0 = type[j0];
        for (; k < total; k += CAPAC_WAVE) {
            const int j2 = nbr_at(k + 2 * CAPAC_WAVE);
            const double x1 = r[j1], y1 = r[N + j1], z1 = r[2 * N + j1], q1 = charge[j1];
            const int t1 = type[j1];
            double dx = x0 - xi, dy = y0 - yi, dz = z0 - zi;
            pbc12(bc, dx, dy, dz);
            double r2 = dist2_exact(dx, dy, dz);
            if (r2 < bc.rc2)                  // rows reach rc + skin
                pair_term(bc, s_tab[ti + t0], qi, q0, dx, dy, dz, r2, o, s_erfcx);
            j1 = j2; x0 = x1; y0 = y1; z0 = z1; q0 = q1; t0 = t1;
        }
        double fx = wave_sum(o.fx), fy = wave_sum(o.fy), fz = wave_sum(o.fz);
        if (lane == 0) {
            if (field) {
                // the external field on the atom's charge, and its energy
                // (force.c:2251-2260), in the same write
                fx += bc.efield[0] * qi; fy += bc.efield[1] * qi; fz += bc.efield[2] * qi;
                o.e9 -= bc.efield[0] * qi * xi + bc.efield[1] * qi * yi + bc.efield[2] * qi * zi;
            }
            if (accumulate) { f[i] += fx; f[N + i] += fy; f[2 * N + i] += fz; }
            else { f[i] = fx; f[N + i] = fy; f[2 * N + i] = fz; }
        }
    }
    const int wave = threadIdx.x / CAPAC_WAVE;
    double vals[12] = {o.e7, o.e8, o.v[0], o.v[1], o.v[2], o.v[1], o.v[3],
                       o.v[4], o.v[2], o.v[4], o.v[5], o.e9};
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        double s = wave_sum(vals[c]);
        if (lane == 0) s_red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double s = 0.0;
        if (threadIdx.x < 12)
            for (int w = 0; w < kWavesPerBlock; ++w) s += s_red[w][threadIdx.x];
        // each ordered pair was visited from both atoms (the field energy,
        // slot 11, once per atom)
        partial[blockIdx.x * 16 + threadIdx.x] = threadIdx.x == 11 ? s : 0.5 * s;
    }
}

}  // namespace

int pair_blocks(int N) {
    int nb = (N + kWavesPerBlock - 1) / kWavesPerBlock;
    return nb < 2048 ? nb : 2048;
}

// writes (or adds to) f; red[0..11) = pot7, pot8, vir[9], and with `field`
// the external field's force on every charge added in the same write and
// red[11] = its energy pot9 (force.c:2251-2260); max_blocks > 0 caps the
// grid (the kernel strides over atoms), leaving compute units to work that
// runs beside it
int pair_nonbonded(capac_ctx* c, double* f_accumulate, int max_blocks, int field) {
    int nblk = pair_blocks(c->bc.N);
    if (max_blocks > 0 && nblk > max_blocks) nblk = max_blocks;
    double* f = f_accumulate ? f_accumulate : c->f;
    k_pair<<<nblk, 256, 0, c->stream>>>(c->r, c->charge, c->type, c->pair, c->bc, c->n_front,
                                        c->n_back, c->nbr, c->nb_stride, f,
                                        f_accumulate ? 1 : 0, field, partial_seg(c, SEG_PAIR));
    CHECK_HIP(hipGetLastError());
    return reduce_partials(c, SEG_PAIR, nblk, 12, 0);
}
