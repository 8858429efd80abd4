// Bonded terms and virtual sites on the device: the reference's mpi_bonded
// (force.c:1694-1904) with compute_bond_1 / compute_bond_3 (force.c:822-909),
// compute_nonbonded for 1-4 pairs (force.c:710-816), compute_angle_1 /
// compute_angle_5 (force.c:956-1052) and compute_dihedral_3 (force.c:1100-1230;
// periodic dihedrals arrive converted to its coefficients), and build_vsites /
// mpi_vsites with compute_vsite_4 (force.c:1060-1090, 1582-1692).
//
// One thread per molecule walks the molecule's terms in param.txt order and
// adds each term's forces to its atoms in the reference's order.  A term only
// touches atoms of its own molecule, so no two threads write one atom, and a
// molecule's bonded forces are the reference's operation for operation
// (contraction off; the transcendental functions are the device library's).
// Energies and the virial are summed per thread and reduced in fixed order.
// Molecules are few and their chains short: one-wavefront workgroups spread
// them over the SIMDs.
#include "capac_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kBondedThreads = 64;
constexpr double kPi = 3.14159265358979323846;   // M_PI

struct V3 {
    double x, y, z;
};

// pbc (force.c:104-109)
__device__ __forceinline__ double pbc(double dx, double L) {
    if (dx < -0.5 * L) dx += L; else if (dx > 0.5 * L) dx -= L;
    return dx;
}

__device__ __forceinline__ V3 load3(const double* __restrict__ r, int N, int i) {
    return V3{r[i], r[N + i], r[2 * N + i]};
}

__device__ __forceinline__ void add3(double* __restrict__ f, int N, int i, const V3& v) {
    f[i] += v.x; f[N + i] += v.y; f[2 * N + i] += v.z;
}

__device__ __forceinline__ V3 pbc_diff(const V3& a, const V3& b, const BoxConst& bc) {
    return V3{pbc(a.x - b.x, bc.box[0]), pbc(a.y - b.y, bc.box[1]), pbc(a.z - b.z, bc.box[2])};
}

// dist_2, inner_product (force.c:41-44, 928-933)
__device__ __forceinline__ double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// outer_product (force.c:940-949)
__device__ __forceinline__ V3 cross(const V3& a, const V3& b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// scalar_x_vector (force.c:915-922)
__device__ __forceinline__ V3 scale(double s, const V3& v) { return V3{s * v.x, s * v.y, s * v.z}; }

// virial_rvec_fvec_full (force.c:84-100): w[a][b] -= r_a f_b
__device__ __forceinline__ void virial_full(const V3& r, const V3& f, double w[9]) {
    w[0] -= r.x * f.x; w[1] -= r.x * f.y; w[2] -= r.x * f.z;
    w[3] -= r.y * f.x; w[4] -= r.y * f.y; w[5] -= r.y * f.z;
    w[6] -= r.z * f.x; w[7] -= r.z * f.y; w[8] -= r.z * f.z;
}

// per-thread energies: pot[2] bonds (and Urey-Bradley), pot[3] angles,
// pot[4] dihedrals, pot[7] and pot[8] of the 1-4 pairs
struct BondedSums {
    double e2, e3, e4, e7, e8, w[9];
};

// compute_bond_1 (force.c:822-861): returns f_j (f_i = -f_j)
__device__ __forceinline__ V3 bond_harmonic(const V3& ri, const V3& rj, double b0, double kb, const BoxConst& bc,
                                            BondedSums& s) {
    const V3 r = pbc_diff(rj, ri, bc);
    const double rij = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    const double fij = kb * (b0 - rij) / rij;
    const V3 fj = scale(fij, r);
    s.e2 += 0.5 * kb * (b0 - rij) * (b0 - rij);
    s.w[0] += -r.x * fj.x; s.w[1] += -r.x * fj.y; s.w[2] += -r.x * fj.z;
    s.w[3] += -r.y * fj.x; s.w[4] += -r.y * fj.y; s.w[5] += -r.y * fj.z;
    s.w[6] += -r.z * fj.x; s.w[7] += -r.z * fj.y; s.w[8] += -r.z * fj.z;
    return fj;
}

// compute_bond_3 (force.c:867-909), Morse bond: returns f_j
__device__ __forceinline__ V3 bond_morse(const V3& ri, const V3& rj, double b0, double D, double beta,
                                         const BoxConst& bc, BondedSums& s) {
    const V3 r = pbc_diff(rj, ri, bc);
    const double rij = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    const double e = exp(-beta * (rij - b0));
    const double fij = -2.0 * D * beta * e * (1.0 - e) / rij;
    const V3 fj = scale(fij, r);
    s.e2 += D * pow(1.0 - e, 2.0);
    s.w[0] += -r.x * fj.x; s.w[1] += -r.x * fj.y; s.w[2] += -r.x * fj.z;
    s.w[3] += -r.y * fj.x; s.w[4] += -r.y * fj.y; s.w[5] += -r.y * fj.z;
    s.w[6] += -r.z * fj.x; s.w[7] += -r.z * fj.y; s.w[8] += -r.z * fj.z;
    return fj;
}

// a 1-4 pair, compute_nonbonded with is_14_pair = 1 (force.c:710-816): plain
// LJ and Coulomb with their scale factors, no cut-off; returns f_j
__device__ __forceinline__ V3 pair14(const V3& ri, const V3& rj, const double* __restrict__ pp, const BoxConst& bc,
                                     BondedSums& s) {
    const V3 r = pbc_diff(rj, ri, bc);
    const double rij2 = r.x * r.x + r.y * r.y + r.z * r.z;
    const double rij = sqrt(rij2);
    const double c6 = pp[0], c12 = pp[1], qi = pp[2], qj = pp[3], sLJ = pp[4], sQQ = pp[5];
    V3 lj = {0.0, 0.0, 0.0}, qq = {0.0, 0.0, 0.0};
    if (c6 != 0.0 || c12 != 0.0) {
        const double rij6 = rij2 * rij2 * rij2;
        const double rij12 = rij6 * rij6;
        const double fij = (c12 * 2.0 / rij12 - c6 / rij6) * 6.0 / rij2 * sLJ;
        s.e8 += (c12 / rij12 - c6 / rij6) * sLJ;
        lj = scale(fij, r);
    }
    if (qi != 0.0 && qj != 0.0) {
        const double fij = sQQ * K_FQQ * qi * qj / (rij2 * rij);
        s.e7 += sQQ * K_FQQ * qi * qj / rij;
        qq = scale(fij, r);
    }
    // virial_vec_nb (force.c:562-583)
    const V3 f = {lj.x + qq.x, lj.y + qq.y, lj.z + qq.z};
    s.w[0] -= r.x * f.x; s.w[1] -= r.x * f.y; s.w[2] -= r.x * f.z;
    s.w[3] -= r.y * f.x; s.w[4] -= r.y * f.y; s.w[5] -= r.y * f.z;
    s.w[6] -= r.z * f.x; s.w[7] -= r.z * f.y; s.w[8] -= r.z * f.z;
    return f;
}

struct Angle3 {
    V3 fi, fj, fk;
};

// compute_angle_1 (force.c:956-1015)
__device__ __forceinline__ Angle3 angle_harmonic(const V3& ri, const V3& rj, const V3& rk, double a0, double ka,
                                                 const BoxConst& bc, BondedSums& s) {
    const V3 r21 = pbc_diff(ri, rj, bc);
    const double r21_d = sqrt(dot(r21, r21));
    const V3 u21 = scale(1.0 / r21_d, r21);
    const V3 r32 = pbc_diff(rj, rk, bc);
    const double r32_d = sqrt(dot(r32, r32));
    const V3 u32 = scale(1.0 / r32_d, r32);
    const double cos123 = -dot(u21, u32);
    const double theta123 = acos(cos123);
    const double sin123 = sin(theta123);
    const double delta_a = a0 * kPi / 180.0 - theta123;
    const double fij = ka * delta_a;
    Angle3 a;
    a.fi.x = (cos123 * u21.x + u32.x) / (sin123 * r21_d) * fij;
    a.fi.y = (cos123 * u21.y + u32.y) / (sin123 * r21_d) * fij;
    a.fi.z = (cos123 * u21.z + u32.z) / (sin123 * r21_d) * fij;
    a.fk.x = -(cos123 * u32.x + u21.x) / (sin123 * r32_d) * fij;
    a.fk.y = -(cos123 * u32.y + u21.y) / (sin123 * r32_d) * fij;
    a.fk.z = -(cos123 * u32.z + u21.z) / (sin123 * r32_d) * fij;
    a.fj.x = -a.fi.x - a.fk.x;
    a.fj.y = -a.fi.y - a.fk.y;
    a.fj.z = -a.fi.z - a.fk.z;
    s.e3 += 0.5 * ka * delta_a * delta_a;
    virial_full(r21, a.fi, s.w);
    virial_full(V3{-r32.x, -r32.y, -r32.z}, a.fk, s.w);
    return a;
}

struct Dihedral4 {
    V3 fi, fj, fk, fl;
};

// compute_dihedral_3 (force.c:1100-1230), Ryckaert-Bellemans in the polymer
// convention
__device__ __forceinline__ Dihedral4 dihedral_rb(const V3& ri, const V3& rj, const V3& rk, const V3& rl,
                                                 const double* __restrict__ c, const BoxConst& bc,
                                                 BondedSums& s) {
    const V3 r21 = pbc_diff(ri, rj, bc);
    const double r21_d = sqrt(dot(r21, r21));
    const V3 u21 = scale(1.0 / r21_d, r21);
    const V3 r32 = pbc_diff(rj, rk, bc);
    const double r32_d = sqrt(dot(r32, r32));
    const V3 u32 = scale(1.0 / r32_d, r32);
    const V3 r43 = pbc_diff(rk, rl, bc);
    const double r43_d = sqrt(dot(r43, r43));
    const V3 u43 = scale(1.0 / r43_d, r43);
    const double cos123 = -dot(u21, u32);
    const double theta123 = acos(cos123);
    const double sin123 = sin(theta123);
    const double cos234 = -dot(u32, u43);
    const double theta234 = acos(cos234);
    const double sin234 = sin(theta234);
    const double cos_phi = (cos123 * cos234 - dot(u21, u43)) / (sin123 * sin234);
    s.e4 += (c[0] - c[1] * cos_phi + c[2] * pow(cos_phi, 2.0) - c[3] * pow(cos_phi, 3.0) +
             c[4] * pow(cos_phi, 4.0) - c[5] * pow(cos_phi, 5.0));
    const double fij = -(-c[1] + c[2] * 2.0 * cos_phi - c[3] * 3.0 * pow(cos_phi, 2.0) +
                         c[4] * 4.0 * pow(cos_phi, 3.0) - c[5] * 5.0 * pow(cos_phi, 4.0));
    const V3 r_21_32 = cross(u21, u32);
    const V3 r_43_32 = cross(u43, u32);
    const double const_1 = dot(u43, r_21_32) / (r21_d * pow(sin123, 3.0) * sin234);
    const double const_4 = dot(u43, r_21_32) / (r43_d * pow(sin234, 3.0) * sin123);
    const double c123 = r21_d * cos123 / r32_d - 1.0;
    const double b432 = r43_d * cos234 / r32_d;
    Dihedral4 d;
    d.fi = V3{-const_1 * r_21_32.x * fij, -const_1 * r_21_32.y * fij, -const_1 * r_21_32.z * fij};
    d.fl = V3{-const_4 * r_43_32.x * fij, -const_4 * r_43_32.y * fij, -const_4 * r_43_32.z * fij};
    d.fj = V3{c123 * d.fi.x - b432 * d.fl.x, c123 * d.fi.y - b432 * d.fl.y, c123 * d.fi.z - b432 * d.fl.z};
    d.fk = V3{-d.fi.x - d.fj.x - d.fl.x, -d.fi.y - d.fj.y - d.fl.y, -d.fi.z - d.fj.z - d.fl.z};
    // -r (x) f with the atoms' positions as the reference forms them: i as
    // stored, then back along the pbc-treated bond vectors
    V3 rv = ri;
    virial_full(rv, d.fi, s.w);
    rv.x -= r21.x; rv.y -= r21.y; rv.z -= r21.z;
    virial_full(rv, d.fj, s.w);
    rv.x -= r32.x; rv.y -= r32.y; rv.z -= r32.z;
    virial_full(rv, d.fk, s.w);
    rv.x -= r43.x; rv.y -= r43.y; rv.z -= r43.z;
    virial_full(rv, d.fl, s.w);
    return d;
}

__global__ void __launch_bounds__(kBondedThreads)
k_bonded(int n_mols, BoxConst bc, const int* __restrict__ mol_terms, const double* __restrict__ r,
         const int* __restrict__ bond_atoms, const int* __restrict__ bond_funct, const double* __restrict__ bond_par,
         const int* __restrict__ pair14_atoms, const double* __restrict__ pair14_par,
         const int* __restrict__ angle_atoms, const int* __restrict__ angle_funct,
         const double* __restrict__ angle_par, const int* __restrict__ dihedral_atoms,
         const double* __restrict__ dihedral_par, double* __restrict__ f, double* __restrict__ partial) {
    __shared__ double s_red[kBondedThreads / 64][14];
    const int N = bc.N;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    BondedSums s = {};
    if (m < n_mols) {
        const int M1 = n_mols + 1;
        for (int t = mol_terms[m]; t < mol_terms[m + 1]; ++t) {
            const int i = bond_atoms[2 * t], j = bond_atoms[2 * t + 1];
            const double* bp = bond_par + 3 * t;
            const V3 ri = load3(r, N, i), rj = load3(r, N, j);
            const V3 fj = bond_funct[t] == 1 ? bond_harmonic(ri, rj, bp[0], bp[1], bc, s)
                                             : bond_morse(ri, rj, bp[0], bp[1], bp[2], bc, s);
            add3(f, N, i, V3{-fj.x, -fj.y, -fj.z});
            add3(f, N, j, fj);
        }
        for (int t = mol_terms[M1 + m]; t < mol_terms[M1 + m + 1]; ++t) {
            const int i = pair14_atoms[2 * t], j = pair14_atoms[2 * t + 1];
            const V3 fj = pair14(load3(r, N, i), load3(r, N, j), pair14_par + 6 * t, bc, s);
            // f_i takes the vdW and Coulomb parts' negatives, summed (force.c:1785-1787)
            add3(f, N, i, V3{-fj.x, -fj.y, -fj.z});
            add3(f, N, j, fj);
        }
        for (int t = mol_terms[2 * M1 + m]; t < mol_terms[2 * M1 + m + 1]; ++t) {
            const int i = angle_atoms[3 * t], j = angle_atoms[3 * t + 1], k = angle_atoms[3 * t + 2];
            const double* ap = angle_par + 4 * t;
            const V3 ri = load3(r, N, i), rj = load3(r, N, j), rk = load3(r, N, k);
            Angle3 a = angle_harmonic(ri, rj, rk, ap[0], ap[1], bc, s);
            if (angle_funct[t] == 5 && ap[3] > 0.0) {
                // Urey-Bradley: the i-k harmonic bond (force.c:1021-1052)
                const V3 fk = bond_harmonic(ri, rk, ap[2], ap[3], bc, s);
                a.fi.x += -fk.x; a.fi.y += -fk.y; a.fi.z += -fk.z;
                a.fk.x += fk.x; a.fk.y += fk.y; a.fk.z += fk.z;
            }
            add3(f, N, i, a.fi);
            add3(f, N, j, a.fj);
            add3(f, N, k, a.fk);
        }
        for (int t = mol_terms[3 * M1 + m]; t < mol_terms[3 * M1 + m + 1]; ++t) {
            const int* q = dihedral_atoms + 4 * t;
            const Dihedral4 d = dihedral_rb(load3(r, N, q[0]), load3(r, N, q[1]), load3(r, N, q[2]),
                                            load3(r, N, q[3]), dihedral_par + 6 * t, bc, s);
            add3(f, N, q[0], d.fi);
            add3(f, N, q[1], d.fj);
            add3(f, N, q[2], d.fk);
            add3(f, N, q[3], d.fl);
        }
    }
    double vals[14] = {s.e2, s.e3, s.e4, s.e7, s.e8, s.w[0], s.w[1], s.w[2], s.w[3], s.w[4],
                       s.w[5], s.w[6], s.w[7], s.w[8]};
#pragma unroll
    for (int c = 0; c < 14; ++c) {
        double v = wave_sum(vals[c]);
        if (lane == 0) s_red[wave][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double v = 0.0;
        if (threadIdx.x < 14)
            for (int w = 0; w < kBondedThreads / 64; ++w) v += s_red[w][threadIdx.x];
        partial[blockIdx.x * 16 + threadIdx.x] = v;
    }
}

// build_vsites (force.c:1582-1629): one thread per molecule, its sites in order
__global__ void __launch_bounds__(kBondedThreads)
k_vsite_build(int n_mols, BoxConst bc, const int* __restrict__ range, const int* __restrict__ atoms,
              const double* __restrict__ par, double* __restrict__ r) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_mols) return;
    const int N = bc.N;
    for (int t = range[m]; t < range[m + 1]; ++t) {
        const int i = atoms[4 * t], j = atoms[4 * t + 1], k = atoms[4 * t + 2], si = atoms[4 * t + 3];
        const double a = par[3 * t], b = par[3 * t + 1], c = par[3 * t + 2];
        const V3 ri = load3(r, N, i);
        const V3 rij = pbc_diff(load3(r, N, j), ri, bc);
        const V3 rik = pbc_diff(load3(r, N, k), ri, bc);
        const V3 rijik = cross(rij, rik);
        r[si] = ri.x + a * rij.x + b * rik.x + c * rijik.x;
        r[N + si] = ri.y + a * rij.y + b * rik.y + c * rijik.y;
        r[2 * N + si] = ri.z + a * rij.z + b * rik.z + c * rijik.z;
    }
}

// mpi_vsites with compute_vsite_4 (force.c:1060-1090, 1636-1692)
__global__ void __launch_bounds__(kBondedThreads)
k_vsite_spread(int n_mols, BoxConst bc, const int* __restrict__ range, const int* __restrict__ atoms,
               const double* __restrict__ par, const double* __restrict__ r, double* __restrict__ f) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_mols) return;
    const int N = bc.N;
    for (int t = range[m]; t < range[m + 1]; ++t) {
        const int i = atoms[4 * t], j = atoms[4 * t + 1], k = atoms[4 * t + 2], si = atoms[4 * t + 3];
        const double a = par[3 * t], b = par[3 * t + 1], c = par[3 * t + 2];
        const V3 ri = load3(r, N, i);
        const V3 dij = pbc_diff(load3(r, N, j), ri, bc);
        const V3 dik = pbc_diff(load3(r, N, k), ri, bc);
        const double fxs = f[si], fys = f[N + si], fzs = f[2 * N + si];
        V3 fj, fk, fi;
        fj.x = a * fxs - c * dik.z * fys + c * dik.y * fzs;
        fj.y = c * dik.z * fxs + a * fys - c * dik.x * fzs;
        fj.z = -c * dik.y * fxs + c * dik.x * fys + a * fzs;
        fk.x = b * fxs + c * dij.z * fys - c * dij.y * fzs;
        fk.y = -c * dij.z * fxs + b * fys + c * dij.x * fzs;
        fk.z = c * dij.y * fxs - c * dij.x * fys + b * fzs;
        fi.x = fxs - fj.x - fk.x;
        fi.y = fys - fj.y - fk.y;
        fi.z = fzs - fj.z - fk.z;
        add3(f, N, i, fi);
        add3(f, N, j, fj);
        add3(f, N, k, fk);
        f[si] = 0.0;
        f[N + si] = 0.0;
        f[2 * N + si] = 0.0;
    }
}

inline unsigned mol_blocks(int n_mols) { return (unsigned)((n_mols + kBondedThreads - 1) / kBondedThreads); }

}  // namespace

bool has_bonded(const capac_ctx* c) { return c->n_bonds + c->n_pairs14 + c->n_angles + c->n_dihedrals > 0; }

int bonded_blocks(int n_mols) { return (int)mol_blocks(n_mols); }

// adds the bonded forces at the resident positions to c->f;
// red[112..126) = pot2, pot3, pot4, pot7, pot8, vir[9]
int bonded_forces(capac_ctx* c) {
    if (!has_bonded(c)) return 0;
    const int nm = c->bc.n_mols;
    const unsigned nblk = mol_blocks(nm);
    k_bonded<<<nblk, kBondedThreads, 0, c->stream>>>(nm, c->bc, c->mol_terms, c->r, c->bond_atoms, c->bond_funct,
                                                      c->bond_par, c->pair14_atoms, c->pair14_par, c->angle_atoms,
                                                      c->angle_funct, c->angle_par, c->dihedral_atoms,
                                                      c->dihedral_par, c->f, partial_seg(c, SEG_BONDED));
    CHECK_HIP(hipGetLastError());
    return reduce_energies(c, SEG_BONDED, (int)nblk, 14, 112);
}

int vsite_build(capac_ctx* c) {
    if (c->n_vsites == 0) return 0;
    const int nm = c->bc.n_mols;
    k_vsite_build<<<mol_blocks(nm), kBondedThreads, 0, c->stream>>>(nm, c->bc, c->mol_terms + 4 * (nm + 1),
                                                                    c->vsite_atoms, c->vsite_par, c->r);
    CHECK_HIP(hipGetLastError());
    return 0;
}

int vsite_spread(capac_ctx* c, double* f) {
    if (c->n_vsites == 0) return 0;
    const int nm = c->bc.n_mols;
    k_vsite_spread<<<mol_blocks(nm), kBondedThreads, 0, c->stream>>>(nm, c->bc, c->mol_terms + 4 * (nm + 1),
                                                                     c->vsite_atoms, c->vsite_par, c->r, f);
    CHECK_HIP(hipGetLastError());
    return 0;
}
