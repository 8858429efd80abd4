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
//
// Two kernels.  k_pair takes every pair type through the general term.
// k_pair_water serves the decks whose non-metal atom types meet only in
// funct 1 pairs under the damped-shifted-force Coulomb and shifted-force LJ
// (use_coulomb = use_vdw = 1) with alpha rc < 6 -- the water of the CPIM
// decks: the front segment of a non-metal atom (its non-metal neighbours,
// ~95% of all row entries) runs a term with no type dispatch and no
// branches: every pair takes the LJ terms (pairs without LJ have c6 = c12 = 0
// and zero shift constants, so they add exact zeros), every pair takes the
// Coulomb terms (a zero charge product adds zeros), exp(-alpha^2 r^2) comes
// from a polynomial whose coefficients sit in scalar registers, and positions
// and charges are read as two 16-byte loads from a packed (x, y, z, q) copy.
// Metal atoms' rows and the back segments (metal neighbours) keep the general
// term, in a second launch, unless every pair type of the deck (the metal's
// included) is funct 1: then the water kernel takes every row.  Results
// differ from k_pair's by rounding only.
//
// An MD step whose energies nobody reads (no save step, no synchronisation)
// runs the water kernel without its energies and virial: the forces are the
// same bits, and the reduction launch is left out.
#include "capac_internal.h"
#include "erfc_fast.h"
#include "exp_poly.h"

namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kPairRestBlocks = 512;    // workgroups of the water kernel's second launch

// what a lane accumulates: the force on the current atom i, and (over all
// of the lane's pairs) the two energies and the virial, symmetric for central
// pair forces, so six components: xx, xy, xz, yy, yz, zz
struct PairOut {
    double fx, fy, fz, e7, e8, e9, v[6];
};

// The reference's expressions with each division by r, r^2, r^6 or the erf
// width replaced by a product with 1/r from one reciprocal square root per
// pair, and r = r^2 (1/r); the shifted-force LJ constants of the pair type
// come from the table (results move by rounding only).  Poly: the
// exponentials from exp_poly and the erf-damped vdW's erf from the erfcx
// table (the water kernel's second launch, which has registers to spare);
// otherwise the library exp and erf (k_pair, which would spill).
template <bool Poly>
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
                double ex = Poly ? exp_poly<false>(-bc.alpha * bc.alpha * r2) : exp(-bc.alpha * bc.alpha * r2);
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
            // (erf and the Gaussian from one exponential: erf = 1 - e^{-x^2}
            // erfcx(x) from the erfcx table for x >= 0.5)
            double rR = rij * pp.inv_width;
            const double ex = Poly ? exp_poly<false>(-rR * rR) : exp(-rR * rR);
            double erf_rR = Poly ? erf_from_exp(rR, ex, erfcx_coef) : erf(rR);
            o.e8 += (cr12 - cr6) * erf_rR;
            flj = ((12.0 * cr12 - 6.0 * cr6) * inv_r * erf_rR -
                   (cr12 - cr6) * 2.0 * K_INV_SQRT_PI * ex * pp.inv_width) * inv_r;
        }
    } else if (pp.funct == 2) {
        // Buckingham (force.c:674-703): A e^{-B r} - C6 / r^6, no Coulomb term
        const double inv_r6 = inv_r2 * inv_r2 * inv_r2;
        const double e = Poly ? exp_poly<false>(-pp.b * rij) : exp(-pp.b * rij);
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
                pair_term<false>(bc, s_tab[ti + t0], qi, q0, dx, dy, dz, r2, o, s_erfcx);
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

// ---------------------------------------------------------------- water path

// the constants of one pair type, as the water term reads them from LDS
// (64 bytes: the entry's offset is a shift of the type index); erf4 = 1 for
// a funct-4 type (erf-damped vdW, 1/width in inv_width)
struct PairLJ {
    double c6, c12, lj_fc, lj_fc_rc, lj_ec, inv_width, erf4, pad;
};

// erf-damped vdW (force.c:602-628, funct 4; pair_term's expressions) without
// a test of c6 and c12 (zeros give zeros); erf from its Maclaurin series
// below 1/2, from the Gaussian and the erfcx table up to 6, 1 beyond.
// Returns the force over r.
template <bool Energy>
__device__ __forceinline__ double water_erf_vdw(const PairLJ& lj, double inv_r, double inv_r2, double rij,
                                                PairOut& o, const double* __restrict__ erfcx_coef) {
    const double inv_r6 = inv_r2 * inv_r2 * inv_r2;
    const double cr12 = lj.c12 * (inv_r6 * inv_r6), cr6 = lj.c6 * inv_r6;
    const double rR = rij * lj.inv_width;
    const double ex = exp_poly<false>(-rR * rR);
    double erf_rR;
    if (rR < 0.5) {
        erf_rR = erf_small(rR);
    } else if (rR < kErfcxHi) {
        const double y = 4.0 * rR;
        const double* c = erfcx_coef + (int)y * kErfcxTerms;
        const double t = fma(y - floor(y), 2.0, -1.0);
        double p = c[kErfcxTerms - 1];
#pragma unroll
        for (int m = kErfcxTerms - 2; m >= 0; --m) p = fma(p, t, c[m]);
        erf_rR = 1.0 - ex * p;
    } else {
        erf_rR = 1.0;
    }
    if (Energy) o.e8 += (cr12 - cr6) * erf_rR;
    return ((12.0 * cr12 - 6.0 * cr6) * inv_r * erf_rR - (cr12 - cr6) * 2.0 * K_INV_SQRT_PI * ex * lj.inv_width) *
           inv_r;
}

// funct 1 under use_coulomb = use_vdw = 1 (pair_term's expressions), for
// every pair without a test of the type or the charges; kqi = K_FQQ q_i,
// na2 = -alpha^2.  lj_row false (the same on every lane): atom i's type has
// no LJ with any type it meets here, so the LJ terms (exact zeros) are left
// out.  A funct-4 pair type (lj.erf4) takes water_erf_vdw instead (no
// Coulomb term, as the reference's compute_erf_vdw).
template <bool Energy, bool Erf4>
__device__ __forceinline__ void water_term(bool lj_row, const BoxConst& bc, const PairLJ& lj, double kqi, double qj,
                                           double na2, double dx, double dy, double dz, double r2, PairOut& o,
                                           const double* __restrict__ erfcx_coef) {
    const double inv_r = rsqrt_pos(r2);
    const double rij = r2 * inv_r;
    const double inv_r2 = inv_r * inv_r;
    if (Erf4 && lj.erf4 != 0.0) {
        const double fs = water_erf_vdw<Energy>(lj, inv_r, inv_r2, rij, o, erfcx_coef);
        o.fx = fma(-fs, dx, o.fx); o.fy = fma(-fs, dy, o.fy); o.fz = fma(-fs, dz, o.fz);
        if (Energy) {
            const double fjx = fs * dx, fjy = fs * dy, fjz = fs * dz;
            o.v[0] -= dx * fjx; o.v[1] -= dx * fjy; o.v[2] -= dx * fjz;
            o.v[3] -= dy * fjy; o.v[4] -= dy * fjz; o.v[5] -= dz * fjz;
        }
        return;
    }
    // shifted-force LJ (force.c:748-757)
    double flj = 0.0;
    if (lj_row) {
        const double inv_r6 = inv_r2 * inv_r2 * inv_r2;
        const double cr12 = lj.c12 * (inv_r6 * inv_r6), cr6 = lj.c6 * inv_r6;
        const double f_r = 12.0 * cr12 - 6.0 * cr6;
        flj = (f_r - lj.lj_fc) * inv_r2;
        if (Energy) o.e8 += (cr12 - cr6) + (rij - bc.rc) * lj.lj_fc_rc - lj.lj_ec;
    }
    // damped shifted force (force.c:787-795); alpha r < 6, so erfc comes from
    // the erfcx table on its whole range: erfc = exp(-x^2) erfcx(x)
    const double qq = kqi * qj;
    const double ex = exp_poly<true>(na2 * r2);
    const double y = 4.0 * (bc.alpha * rij);
    const int piece = (int)y;
    const double t = fma(y - floor(y), 2.0, -1.0);
    const double* c = erfcx_coef + piece * kErfcxTerms;
    double p = c[kErfcxTerms - 1];
#pragma unroll
    for (int m = kErfcxTerms - 2; m >= 0; --m) p = fma(p, t, c[m]);
    const double erfc_a = ex * p;
    const double fqq = qq * ((erfc_a * inv_r2 + bc.a2_sqrtPI * ex * inv_r) - bc.wolfConst) * inv_r;
    if (Energy) o.e7 += qq * (erfc_a * inv_r - bc.erfc_arCut + bc.wolfConst * (rij - bc.rc));
    const double fs = flj + fqq;
    // explicit FMAs: the force takes the same instructions with or without
    // the energies beside it
    o.fx = fma(-fs, dx, o.fx); o.fy = fma(-fs, dy, o.fy); o.fz = fma(-fs, dz, o.fz);
    if (Energy) {
        const double fjx = fs * dx, fjy = fs * dy, fjz = fs * dz;
        o.v[0] -= dx * fjx; o.v[1] -= dx * fjy; o.v[2] -= dx * fjz;
        o.v[3] -= dy * fjy; o.v[4] -= dy * fjz; o.v[5] -= dz * fjz;
    }
}

// atom i's row with the water term: the nf entries of its front segment
// (row), then cnt - nf of its back segment (back).  Software pipeline three rounds deep: while round s is evaluated, the
// positions and charges of rounds s + 1 and s + 2 and the neighbour index of
// round s + 3 are in flight (beside the solve the kernel has two wavefronts
// per SIMD, too few to hide a gather's latency behind one round of work:
// with two rounds in flight the Au2057 launch beside the solve took 0.79 ms
// per MD step, with three 0.69 ms).  Unrolled three times, so that each
// round's data keep their registers.
template <bool Energy, bool Erf4>
__device__ __forceinline__ void water_segment(bool lj_row, const double2* __restrict__ xyzq,
                                              const int* __restrict__ type, const int* __restrict__ row, int nf,
                                              const int* __restrict__ back, int cnt,
                                              int i, int lane, const BoxConst& bc, const PairLJ* s_lj, int ti,
                                              double xi, double yi, double zi, double kqi, double na2, PairOut& o,
                                              const double* s_erfcx) {
    // past the row: the atom itself (a valid address for the look-ahead loads)
    auto nbr_at = [&](int kk) { return kk < nf ? row[kk] : (kk < cnt ? back[kk - nf] : i); };
    auto round = [&](const double2& a, const double2& b, int t) {
        double dx = a.x - xi, dy = a.y - yi, dz = b.x - zi;
        pbc12(bc, dx, dy, dz);
        const double r2 = dist2_exact(dx, dy, dz);
        if (r2 < bc.rc2)                   // rows reach rc + skin
            water_term<Energy, Erf4>(lj_row, bc, s_lj[ti + t], kqi, b.y, na2, dx, dy, dz, r2, o, s_erfcx);
    };
    constexpr int W = CAPAC_WAVE;
    int k = lane;
    int i0 = nbr_at(k), i1 = nbr_at(k + W), i2 = nbr_at(k + 2 * W);
    double2 a0 = xyzq[2 * i0], b0 = xyzq[2 * i0 + 1];
    int t0 = type[i0];
    double2 a1 = xyzq[2 * i1], b1 = xyzq[2 * i1 + 1];
    int t1 = type[i1];
    double2 a2, b2;
    int t2;
    i0 = nbr_at(k + 3 * W);
    while (k < cnt) {
        a2 = xyzq[2 * i2]; b2 = xyzq[2 * i2 + 1]; t2 = type[i2];
        i1 = nbr_at(k + 4 * W);
        round(a0, b0, t0);
        k += W;
        if (k >= cnt) break;
        a0 = xyzq[2 * i0]; b0 = xyzq[2 * i0 + 1]; t0 = type[i0];
        i2 = nbr_at(k + 4 * W);
        round(a1, b1, t1);
        k += W;
        if (k >= cnt) break;
        a1 = xyzq[2 * i1]; b1 = xyzq[2 * i1 + 1]; t1 = type[i1];
        i0 = nbr_at(k + 4 * W);
        round(a2, b2, t2);
        k += W;
    }
}

// the 12 partial sums of a workgroup (pot7, pot8, the virial, the field
// energy), written to partial[slot * 16 ..]; each ordered pair was visited
// from both atoms (the field energy, slot 11, once per atom)
__device__ __forceinline__ void pair_block_partials(const PairOut& o, double (*s_red)[12], int lane,
                                                    double* __restrict__ partial, int slot) {
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
        partial[slot * 16 + threadIdx.x] = threadIdx.x == 11 ? s : 0.5 * s;
    }
}

// Water kernel, one wavefront per atom; writes (or adds to) the forces of
// the atoms it takes.  all_rows = 0: the front segments of the non-metal
// atoms, the first of two launches.  all_rows = 1 (every pair type funct 1):
// the whole rows of all atoms, the metal atoms' back segments (their metal
// neighbours) left out when metal_inert says they add exact zeros.  Energy =
// false: forces only (no energies, virial or partial sums).
template <bool Energy>
__global__ void __launch_bounds__(256, 3)
k_pair_water(const double2* __restrict__ xyzq, const int* __restrict__ type, const PairParam* __restrict__ table,
             BoxConst bc, const int* __restrict__ n_front, const int* __restrict__ n_back,
             const int* __restrict__ nbr, int stride, double* __restrict__ f, int accumulate, int field,
             double* __restrict__ partial, int lj_types, int all_rows, int metal_inert) {
    __shared__ PairLJ s_lj[64];
    __shared__ double s_red[kWavesPerBlock][12];
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    const int N = bc.N, nt = bc.n_types;
    for (int k = threadIdx.x; k < nt * nt && k < 64; k += blockDim.x) {
        const PairParam& pp = table[k];
        const bool e4 = pp.funct == 4;
        s_lj[k] = PairLJ{pp.c6, pp.c12, pp.lj_fc, pp.lj_fc_rc, pp.lj_ec, e4 ? pp.inv_width : 0.0, e4 ? 1.0 : 0.0, 0.0};
    }
    load_erfcx_table(s_erfcx);
    __syncthreads();
    const int lane = threadIdx.x & (CAPAC_WAVE - 1);
    const double na2 = -bc.alpha * bc.alpha;
    const int first = __builtin_amdgcn_readfirstlane((all_rows ? 0 : bc.Nm) + blockIdx.x * kWavesPerBlock +
                                                     threadIdx.x / CAPAC_WAVE);
    PairOut o = {};
    for (int i = first; i < N; i += gridDim.x * kWavesPerBlock) {
        const double2 ai = xyzq[2 * i], bi = xyzq[2 * i + 1];
        const double xi = ai.x, yi = ai.y, zi = bi.x, qi = bi.y;
        const int ti = type[i] * nt;
        o.fx = 0.0; o.fy = 0.0; o.fz = 0.0;
        const bool lj_row = (lj_types >> type[i]) & 1;
        const int* row = nbr + (size_t)i * stride;
        const int nf = n_front[i];
        const int nb = all_rows && !(i < bc.Nm && metal_inert) ? n_back[i] : 0;
        // a non-metal atom without metal neighbours meets funct-1 types only:
        // its loop leaves out the funct-4 test
        if (i >= bc.Nm && nb == 0)
            water_segment<Energy, false>(lj_row, xyzq, type, row, nf, row, nf, i, lane, bc, s_lj, ti, xi, yi, zi,
                                         K_FQQ * qi, na2, o, s_erfcx);
        else
            water_segment<Energy, true>(lj_row, xyzq, type, row, nf, row + stride - nb, nf + nb, i, lane, bc, s_lj,
                                        ti, xi, yi, zi, K_FQQ * qi, na2, o, s_erfcx);
        double fx = wave_sum(o.fx), fy = wave_sum(o.fy), fz = wave_sum(o.fz);
        if (lane == 0) {
            if (field) {
                // the external field on the atom's charge, and its energy
                // (force.c:2251-2260), in the same write
                fx += bc.efield[0] * qi; fy += bc.efield[1] * qi; fz += bc.efield[2] * qi;
                if (Energy) o.e9 -= bc.efield[0] * qi * xi + bc.efield[1] * qi * yi + bc.efield[2] * qi * zi;
            }
            if (accumulate) { f[i] += fx; f[N + i] += fy; f[2 * N + i] += fz; }
            else { f[i] = fx; f[N + i] = fy; f[2 * N + i] = fz; }
        }
    }
    if (Energy) pair_block_partials(o, s_red, lane, partial, blockIdx.x);
}

// Water kernel, second launch: what the first leaves to the general term --
// the whole rows of the metal atoms (their forces written, or added to; only
// the front segment when metal-metal pairs add nothing) and
// the back segments (metal neighbours) of the non-metal atoms (added to
// their forces).  Wavefront w of W takes metal atoms w, w + W, ..., then
// the non-metal atoms Nm + w + k W (k = 0, 1, ...) that have a back
// segment: its lanes test 64 of them at once and it takes those found in
// order.  The non-metal atoms near the metal are runs of consecutive
// indices; the stride spreads them evenly over the wavefronts.  The
// assignment is fixed, so each workgroup's sums keep a fixed order.
__global__ void __launch_bounds__(256, 2)
k_pair_rest(const double* __restrict__ r, const double* __restrict__ charge, const int* __restrict__ type,
            const PairParam* __restrict__ table, BoxConst bc, const int* __restrict__ n_front,
            const int* __restrict__ n_back, const int* __restrict__ nbr, int stride, double* __restrict__ f,
            int accumulate, int field, double* __restrict__ partial, int slot0, int metal_inert) {
    __shared__ PairParam s_tab[64];
    __shared__ double s_red[kWavesPerBlock][12];
    __shared__ double s_erfcx[kErfcxPieces * kErfcxTerms];
    const int N = bc.N, Nm = bc.Nm, nt = bc.n_types;
    for (int k = threadIdx.x; k < nt * nt && k < 64; k += blockDim.x) s_tab[k] = table[k];
    load_erfcx_table(s_erfcx);
    __syncthreads();
    const int lane = threadIdx.x & (CAPAC_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / CAPAC_WAVE);
    const int n_waves = gridDim.x * kWavesPerBlock;
    PairOut o = {};
    // atom i's row from entry `skip` on (metal: all of it; others: the back segment)
    auto do_atom = [&](int i, bool metal) {
        const double xi = r[i], yi = r[N + i], zi = r[2 * N + i];
        const double qi = charge[i];
        const int ti = type[i] * nt;
        const int nf = n_front[i], nb = n_back[i];
        const int* row = nbr + (size_t)i * stride;
        // (a metal atom's back segment, its metal neighbours, adds exact
        // zeros when the metal is inert)
        const int total = metal ? (metal_inert ? nf : nf + nb) : nb;
        const int skip = metal ? 0 : nf;
        o.fx = 0.0; o.fy = 0.0; o.fz = 0.0;
        for (int k = lane; k < total; k += CAPAC_WAVE) {
            const int kk = k + skip;
            const int j = kk < nf ? row[kk] : row[stride - nb + (kk - nf)];
            double dx = r[j] - xi, dy = r[N + j] - yi, dz = r[2 * N + j] - zi;
            pbc12(bc, dx, dy, dz);
            const double r2 = dist2_exact(dx, dy, dz);
            if (r2 < bc.rc2) pair_term<true>(bc, s_tab[ti + type[j]], qi, charge[j], dx, dy, dz, r2, o, s_erfcx);
        }
        double fx = wave_sum(o.fx), fy = wave_sum(o.fy), fz = wave_sum(o.fz);
        if (lane == 0) {
            if (metal && field) {
                fx += bc.efield[0] * qi; fy += bc.efield[1] * qi; fz += bc.efield[2] * qi;
                o.e9 -= bc.efield[0] * qi * xi + bc.efield[1] * qi * yi + bc.efield[2] * qi * zi;
            }
            if (accumulate || !metal) { f[i] += fx; f[N + i] += fy; f[2 * N + i] += fz; }
            else { f[i] = fx; f[N + i] = fy; f[2 * N + i] = fz; }
        }
    };
    const int w = blockIdx.x * kWavesPerBlock + wave;
    for (int i = w; i < Nm; i += n_waves) do_atom(i, true);
    for (long base = (long)Nm + w; base < N; base += (long)CAPAC_WAVE * n_waves) {
        const long mine = base + (long)lane * n_waves;
        unsigned long long todo = __ballot(mine < N && n_back[mine] > 0);
        while (todo) {
            const int i = (int)(base + (long)__builtin_ctzll(todo) * n_waves);
            todo &= todo - 1;
            do_atom(i, false);
        }
    }
    pair_block_partials(o, s_red, lane, partial, slot0 + blockIdx.x);
}

// the packed (x, y, z, q) copy the water kernel reads: two double2 per atom
__global__ void __launch_bounds__(256)
k_pack_xyzq(const double* __restrict__ r, const double* __restrict__ charge, int N, double2* __restrict__ xyzq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    xyzq[2 * i] = make_double2(r[i], r[N + i]);
    xyzq[2 * i + 1] = make_double2(r[2 * N + i], charge[i]);
}

}  // namespace

int pair_blocks(int N) {
    int nb = (N + kWavesPerBlock - 1) / kWavesPerBlock;
    return nb < 2048 ? nb : 2048;
}

// partial-sum slots the pair launches may write: the water kernel's second
// launch adds its workgroups after the first's
int pair_partial_blocks(int N) { return pair_blocks(N) + kPairRestBlocks; }

// writes (or adds to) f; red[0..11) = pot7, pot8, vir[9], and with `field`
// the external field's force on every charge added in the same write and
// red[11] = its energy pot9 (force.c:2251-2260); max_blocks > 0 caps the
// grid (the kernel strides over atoms), leaving compute units to work that
// runs beside it
int pair_nonbonded(capac_ctx* c, double* f_accumulate, int max_blocks, int field) {
    int nblk = pair_blocks(c->bc.N);
    if (max_blocks > 0 && nblk > max_blocks) nblk = max_blocks;
    double* f = f_accumulate ? f_accumulate : c->f;
    if (c->pair_water) {
        double2* xyzq = reinterpret_cast<double2*>(c->xyzq);
        double* part = partial_seg(c, SEG_PAIR);
        const int all_rows = c->pair_water == 2;
        k_pack_xyzq<<<(c->bc.N + 255) / 256, 256, 0, c->stream>>>(c->r, c->charge, c->bc.N, xyzq);
        if (c->need_energy)
            k_pair_water<true><<<nblk, 256, 0, c->stream>>>(xyzq, c->type, c->pair, c->bc, c->n_front, c->n_back,
                                                            c->nbr, c->nb_stride, f, f_accumulate ? 1 : 0, field,
                                                            part, c->pair_lj_types, all_rows, c->pair_metal_inert);
        else
            k_pair_water<false><<<nblk, 256, 0, c->stream>>>(xyzq, c->type, c->pair, c->bc, c->n_front, c->n_back,
                                                             c->nbr, c->nb_stride, f, f_accumulate ? 1 : 0, field,
                                                             part, c->pair_lj_types, all_rows, c->pair_metal_inert);
        int nparts = nblk;
        if (!all_rows) {
            k_pair_rest<<<kPairRestBlocks, 256, 0, c->stream>>>(c->r, c->charge, c->type, c->pair, c->bc,
                                                                c->n_front, c->n_back, c->nbr, c->nb_stride, f,
                                                                f_accumulate ? 1 : 0, field, part, nblk,
                                                                c->pair_metal_inert);
            nparts += kPairRestBlocks;
        }
        CHECK_HIP(hipGetLastError());
        return reduce_energies(c, SEG_PAIR, nparts, 12, 0);
    } else {
        k_pair<<<nblk, 256, 0, c->stream>>>(c->r, c->charge, c->type, c->pair, c->bc, c->n_front,
                                            c->n_back, c->nbr, c->nb_stride, f,
                                            f_accumulate ? 1 : 0, field, partial_seg(c, SEG_PAIR));
    }
    CHECK_HIP(hipGetLastError());
    return reduce_energies(c, SEG_PAIR, nblk, 12, 0);
}
