// MD step body on the device (reference mpi-main.c:874-1013): velocity
// Verlet half kicks and drift, RATTLE (force.c:1907-2090, one thread per
// molecule, Gauss-Seidel over its constraints in file order), kinetic energy,
// the Nose-Hoover chain (thermostat.c:117-191, one thread: M=10, Nc=5,
// Suzuki-Yoshida order 5), centre-of-mass removal (thermostat.c:69-108) and
// the molecule wrap (force.c:127-165).  All FP64; reductions are fixed-order
// single-workgroup sums so repeated runs are bit-identical.
//
// md[] layout: [0] eKsum [1] T [2..12) vQ [12..22) aQ [22] scale [23] F_max
//              [24] F_rms [25] qMass
#include "capac_internal.h"

// Integrator arithmetic is evaluated as written, without fused multiply-add,
// as the reference's x86-64 build does, so that RATTLE's tolerance decisions
// see the same numbers.
#pragma clang fp contract(off)

namespace {

constexpr int kRedThreads = 1024;

__device__ double block_sum1024(double v, double* s_w) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) s_w[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += s_w[w];
    return t;
}

// the half kick of atom i (mpi-main.c:944-949); fixed metal atoms stay
__device__ __forceinline__ void kick_atom(int i, int N, int Nm, int fix_metal, double dt_2,
                                          const double* __restrict__ f, const double* __restrict__ inv_mass,
                                          double* __restrict__ v) {
    if (fix_metal == 1 && i < Nm) return;
    double im = inv_mass[i];
    v[i] += dt_2 * f[i] * im;
    v[N + i] += dt_2 * f[N + i] * im;
    v[2 * N + i] += dt_2 * f[2 * N + i] * im;
}

__global__ void k_kick(int N, int Nm, int fix_metal, double dt_2, const double* __restrict__ f,
                       const double* __restrict__ inv_mass, double* __restrict__ v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    kick_atom(i, N, Nm, fix_metal, dt_2, f, inv_mass, v);
}

// First half of a step in one pass: the thermostat's velocity scale (nvt:
// every atom, as k_scale_v), then, for every atom that moves, the half kick
// and the drift of k_kick and k_drift; the same operations in the same order.
// pre_com: first the previous step's last pass (k_scale_com with its scale
// in md[29] and the centre-of-mass velocity in md[26..28]), the same
// operations in the same order.
__global__ void k_scale_kick_drift(int N, int Nm, int fix_metal, int scale, double dt_2, double dt,
                                   const double* __restrict__ md, const double* __restrict__ f,
                                   const double* __restrict__ inv_mass, double* __restrict__ v,
                                   double* __restrict__ r, double* __restrict__ old_r, int pre_com,
                                   const double* __restrict__ mass) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (pre_com) {
        const double s = md[29];
        const bool has_mass = mass[i] > 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t k = (size_t)c * N + i;
            double vk = v[k] * s;
            if (has_mass) vk -= md[26 + c];
            v[k] = vk;
        }
    }
    if (scale) {
        const double s = md[22];
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c * N + i] *= s;
    }
    if (fix_metal == 1 && i < Nm) return;
    double im = inv_mass[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t k = (size_t)c * N + i;
        v[k] += dt_2 * f[k] * im;
        old_r[k] = r[k];
        r[k] += dt * v[k];
    }
}

// pbc (force.c:104-109)
__device__ __forceinline__ double pbc1(double dx, double L) {
    if (dx < -0.5 * L) dx += L; else if (dx > 0.5 * L) dx -= L;
    return dx;
}

// Three-site molecules whose constraints are (0,1), (0,2), (1,2) in file
// order (rigid water) run a copy of the Gauss-Seidel loop with compile-time
// atom indices, so the whole molecule stays in registers; the loop-invariant
// parts leave the loop (Cstr3 below) and positions and velocities are
// identical to the general loop's.
__device__ __forceinline__ bool is_rigid3(int f0, int nat, int c0, int nc, const int* __restrict__ ci,
                                          const int* __restrict__ cj) {
    return nat == 3 && nc == 3 && ci[c0] == f0 && cj[c0] == f0 + 1 && ci[c0 + 1] == f0 &&
           cj[c0 + 1] == f0 + 2 && ci[c0 + 2] == f0 + 1 && cj[c0 + 2] == f0 + 2;
}

struct Site {
    double x, y, z, vx, vy, vz, ox, oy, oz, im;
};

// Rigid three-site molecules: the reference's Gauss-Seidel sweeps with what
// does not change inside the loop computed once per constraint (a, b): the
// old-position vector q (RATTLE_1) or the position vector p (RATTLE_2), the
// factor 2 (1/m_a + 1/m_b) or (1/m_a + 1/m_b) r_ab^2, and the tolerance
// bound, each formed as the reference forms it, so positions and velocities
// stay bitwise the reference's.  (A constraint's tolerance test can flip on a
// last-bit change, and one sweep more or less moves a hydrogen by ~1e-6 nm.)
// Only the RATTLE_2 virial, which feeds nothing back, is deferred: it is
// linear in the multipliers g of a constraint whose p is fixed, so the sweeps
// carry the sum G and the virial is formed once at the end (rounding-level
// differences from the reference's per-sweep sums).
struct Cstr3 {
    double vx, vy, vz;     // q (RATTLE_1) or p (RATTLE_2)
    double rab2, bound, den;
};

__device__ __forceinline__ Cstr3 cstr1_setup(const Site& a, const Site& b, double rab, const BoxConst& bc) {
    const double tol2 = 2.0e-05;
    Cstr3 c;
    c.vx = pbc1(a.ox - b.ox, bc.box[0]);
    c.vy = pbc1(a.oy - b.oy, bc.box[1]);
    c.vz = pbc1(a.oz - b.oz, bc.box[2]);
    c.rab2 = rab * rab;
    c.bound = c.rab2 * tol2;
    c.den = 2.0 * (a.im + b.im);
    return c;
}

// x / d for a fixed divisor d, bitwise the IEEE quotient: with inv_d = RN(1/d)
// and q within an ulp of x/d, RN(q + (x - q d) inv_d) is the correctly
// rounded quotient (Markstein's theorem; x - q d is exact in one fma).  The
// fixed-divisor identity was also checked on 2e8 random quotients by ten
// step sizes on the host.  Three dependent operations instead of a division.
__device__ __forceinline__ double div_fixed(double x, double d, double inv_d) {
    const double q = x * inv_d;
    return fma(fma(-q, d, x), inv_d, q);
}

// one RATTLE_1 correction of constraint (a, b) (force.c:1949-1988); inv_dt = 1/dt
__device__ __forceinline__ bool rattle1_pair(Site& a, Site& b, const Cstr3& c, const BoxConst& bc, double dt,
                                             double inv_dt) {
    double px = pbc1(a.x - b.x, bc.box[0]);
    double py = pbc1(a.y - b.y, bc.box[1]);
    double pz = pbc1(a.z - b.z, bc.box[2]);
    double pab2 = px * px + py * py + pz * pz;
    double diff = c.rab2 - pab2;
    if (!(fabs(diff) > c.bound)) return false;
    double rp = c.vx * px + c.vy * py + c.vz * pz;
    double g = diff / (c.den * rp);
    double dx = c.vx * g, dy = c.vy * g, dz = c.vz * g;
    const double ma = a.im, mb = b.im;
    a.x += ma * dx; a.y += ma * dy; a.z += ma * dz;
    b.x -= mb * dx; b.y -= mb * dy; b.z -= mb * dz;
    dx = div_fixed(dx, dt, inv_dt); dy = div_fixed(dy, dt, inv_dt); dz = div_fixed(dz, dt, inv_dt);
    a.vx += ma * dx; a.vy += ma * dy; a.vz += ma * dz;
    b.vx -= mb * dx; b.vy -= mb * dy; b.vz -= mb * dz;
    return true;
}

__device__ __forceinline__ Cstr3 cstr2_setup(const Site& a, const Site& b, double rab, const BoxConst& bc) {
    Cstr3 c;
    c.vx = pbc1(a.x - b.x, bc.box[0]);
    c.vy = pbc1(a.y - b.y, bc.box[1]);
    c.vz = pbc1(a.z - b.z, bc.box[2]);
    c.rab2 = rab * rab;
    c.bound = 1.0e-05;
    c.den = (a.im + b.im) * c.rab2;
    return c;
}

// one RATTLE_2 correction of constraint (a, b) (force.c:2046-2081)
__device__ __forceinline__ bool rattle2_pair(Site& a, Site& b, const Cstr3& c, double& G) {
    double ux = a.vx - b.vx, uy = a.vy - b.vy, uz = a.vz - b.vz;
    double rv = c.vx * ux + c.vy * uy + c.vz * uz;
    double g = -rv / c.den;
    if (!(fabs(g) > c.bound)) return false;
    double dx = c.vx * g, dy = c.vy * g, dz = c.vz * g;
    const double ma = a.im, mb = b.im;
    a.vx += ma * dx; a.vy += ma * dy; a.vz += ma * dz;
    b.vx -= mb * dx; b.vy -= mb * dy; b.vz -= mb * dz;
    G += g;
    return true;
}

// the constraint virial of (a, b) for the multiplier sum G: w -= p (p G / (dt/2))
__device__ __forceinline__ void rattle2_virial(const Cstr3& c, double G, double half_dt, double w[9]) {
    const double h = G / half_dt;
    const double d[3] = {c.vx * h, c.vy * h, c.vz * h};
    const double pv[3] = {c.vx, c.vy, c.vz};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) w[a * 3 + b] -= pv[a] * d[b];
}

__device__ __forceinline__ void load_site(Site& s, int i, int N, const double* __restrict__ r,
                                          const double* __restrict__ v, const double* __restrict__ old_r,
                                          const double* __restrict__ inv_mass) {
    s.x = r[i]; s.y = r[N + i]; s.z = r[2 * N + i];
    s.vx = v[i]; s.vy = v[N + i]; s.vz = v[2 * N + i];
    if (old_r) { s.ox = old_r[i]; s.oy = old_r[N + i]; s.oz = old_r[2 * N + i]; }
    s.im = inv_mass[i];
}

// RATTLE: one thread per molecule, and each thread a long serial chain of
// FP64 operations; small workgroups spread the chains over more SIMDs, so
// that fewer of them share a SIMD's issue slots
constexpr int kRattleThreads = 64;

// One thread per molecule; Gauss-Seidel sweeps over its constraints in file
// order.  The general loop keeps the reference's arithmetic (no contraction,
// see the file pragma); rigid three-site molecules take the register path.
__global__ void __launch_bounds__(kRattleThreads) k_rattle1(int n_mols, int N, BoxConst bc, double dt, const int* __restrict__ mol_cstr,
                          const int* __restrict__ mol_first, const int* __restrict__ mol_last,
                          const int* __restrict__ ci, const int* __restrict__ cj,
                          const double* __restrict__ cr, const double* __restrict__ inv_mass,
                          double* __restrict__ r, double* __restrict__ v,
                          const double* __restrict__ old_r, int* __restrict__ status) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_mols) return;
    int c0 = mol_cstr[m], c1 = mol_cstr[m + 1];
    if (c0 == c1) return;
    const int f0 = mol_first[m], nat = mol_last[m] - f0 + 1, nc = c1 - c0;
    if (is_rigid3(f0, nat, c0, nc, ci, cj)) {
        Site s0, s1, s2;
        load_site(s0, f0, N, r, v, old_r, inv_mass);
        load_site(s1, f0 + 1, N, r, v, old_r, inv_mass);
        load_site(s2, f0 + 2, N, r, v, old_r, inv_mass);
        const Cstr3 k01 = cstr1_setup(s0, s1, cr[c0], bc), k02 = cstr1_setup(s0, s2, cr[c0 + 1], bc),
                    k12 = cstr1_setup(s1, s2, cr[c0 + 2], bc);
        const double inv_dt = 1.0 / dt;
        int iter = 0, done = 0;
        while (!done) {
            done = 1;
            if (rattle1_pair(s0, s1, k01, bc, dt, inv_dt)) done = 0;
            if (rattle1_pair(s0, s2, k02, bc, dt, inv_dt)) done = 0;
            if (rattle1_pair(s1, s2, k12, bc, dt, inv_dt)) done = 0;
            if (++iter > 1000) break;
        }
        const Site* st[3] = {&s0, &s1, &s2};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int i = f0 + q;
            r[i] = st[q]->x; r[N + i] = st[q]->y; r[2 * N + i] = st[q]->z;
            v[i] = st[q]->vx; v[N + i] = st[q]->vy; v[2 * N + i] = st[q]->vz;
        }
        if (!done) status[0] = 1;
        return;
    }
    const double tol2 = 2.0e-05;
    double *rx = r, *ry = r + N, *rz = r + 2 * N;
    double *vx = v, *vy = v + N, *vz = v + 2 * N;
    int iter = 0, done = 0;
    while (!done) {
        done = 1;
        for (int c = c0; c < c1; ++c) {
            int a = ci[c], b = cj[c];
            double rab = cr[c];
            double px = pbc1(rx[a] - rx[b], bc.box[0]);
            double py = pbc1(ry[a] - ry[b], bc.box[1]);
            double pz = pbc1(rz[a] - rz[b], bc.box[2]);
            double pab2 = px * px + py * py + pz * pz;
            double rab2 = rab * rab;
            double diff = rab2 - pab2;
            if (fabs(diff) > rab2 * tol2) {
                double qx = pbc1(old_r[a] - old_r[b], bc.box[0]);
                double qy = pbc1(old_r[N + a] - old_r[N + b], bc.box[1]);
                double qz = pbc1(old_r[2 * N + a] - old_r[2 * N + b], bc.box[2]);
                double rp = qx * px + qy * py + qz * pz;
                double ma = inv_mass[a], mb = inv_mass[b];
                double g = diff / (2.0 * (ma + mb) * rp);
                double dx = qx * g, dy = qy * g, dz = qz * g;
                rx[a] += ma * dx; ry[a] += ma * dy; rz[a] += ma * dz;
                rx[b] -= mb * dx; ry[b] -= mb * dy; rz[b] -= mb * dz;
                dx /= dt; dy /= dt; dz /= dt;
                vx[a] += ma * dx; vy[a] += ma * dy; vz[a] += ma * dz;
                vx[b] -= mb * dx; vy[b] -= mb * dy; vz[b] -= mb * dz;
                done = 0;
            }
        }
        if (++iter > 1000) break;
    }
    if (!done) status[0] = 1;
}

__global__ void __launch_bounds__(kRattleThreads)
k_rattle2(int n_mols, int N, BoxConst bc, double dt, const int* __restrict__ mol_cstr,
          const int* __restrict__ mol_first, const int* __restrict__ mol_last,
          const int* __restrict__ ci, const int* __restrict__ cj, const double* __restrict__ cr,
          const double* __restrict__ inv_mass, const double* __restrict__ r, double* __restrict__ v,
          int* __restrict__ status, double* __restrict__ partial, int Nm, int fix_metal,
          const double* __restrict__ f) {
    __shared__ double s_red[kRattleThreads / 64][9];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int c0m = m < n_mols ? mol_cstr[m] : 0, c1m = m < n_mols ? mol_cstr[m + 1] : 0;
    const int f0 = m < n_mols ? mol_first[m] : 0;
    const int nat = m < n_mols ? mol_last[m] - f0 + 1 : 0, ncm = c1m - c0m;
    // f non-NULL: the step's half kick of the molecule's atoms first (every
    // atom is in one molecule), the same operations as k_kick
    if (f)
        for (int a = f0; a < f0 + nat; ++a) kick_atom(a, N, Nm, fix_metal, 0.5 * dt, f, inv_mass, v);
    if (m < n_mols && is_rigid3(f0, nat, c0m, ncm, ci, cj)) {
        const double half_dt = 0.5 * dt;
        Site s0, s1, s2;
        load_site(s0, f0, N, r, v, NULL, inv_mass);
        load_site(s1, f0 + 1, N, r, v, NULL, inv_mass);
        load_site(s2, f0 + 2, N, r, v, NULL, inv_mass);
        const Cstr3 k01 = cstr2_setup(s0, s1, cr[c0m], bc), k02 = cstr2_setup(s0, s2, cr[c0m + 1], bc),
                    k12 = cstr2_setup(s1, s2, cr[c0m + 2], bc);
        double g01 = 0.0, g02 = 0.0, g12 = 0.0;
        int iter = 0, done = 0;
        while (!done) {
            done = 1;
            if (rattle2_pair(s0, s1, k01, g01)) done = 0;
            if (rattle2_pair(s0, s2, k02, g02)) done = 0;
            if (rattle2_pair(s1, s2, k12, g12)) done = 0;
            if (++iter > 1000) break;
        }
        rattle2_virial(k01, g01, half_dt, w);
        rattle2_virial(k02, g02, half_dt, w);
        rattle2_virial(k12, g12, half_dt, w);
        const Site* st[3] = {&s0, &s1, &s2};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int i = f0 + q;
            v[i] = st[q]->vx; v[N + i] = st[q]->vy; v[2 * N + i] = st[q]->vz;
        }
        if (!done) status[1] = 1;
    } else if (m < n_mols) {
        int c0 = c0m, c1 = c1m;
        const double tol = 1.0e-05, half_dt = 0.5 * dt;
        const double *rx = r, *ry = r + N, *rz = r + 2 * N;
        double *vx = v, *vy = v + N, *vz = v + 2 * N;
        int iter = 0, done = c0 == c1;
        while (!done) {
            done = 1;
            for (int c = c0; c < c1; ++c) {
                int a = ci[c], b = cj[c];
                double rab = cr[c];
                double ux = vx[a] - vx[b], uy = vy[a] - vy[b], uz = vz[a] - vz[b];
                double px = pbc1(rx[a] - rx[b], bc.box[0]);
                double py = pbc1(ry[a] - ry[b], bc.box[1]);
                double pz = pbc1(rz[a] - rz[b], bc.box[2]);
                double rv = px * ux + py * uy + pz * uz;
                double ma = inv_mass[a], mb = inv_mass[b];
                double g = -rv / ((ma + mb) * (rab * rab));
                if (fabs(g) > tol) {
                    double dx = px * g, dy = py * g, dz = pz * g;
                    vx[a] += ma * dx; vy[a] += ma * dy; vz[a] += ma * dz;
                    vx[b] -= mb * dx; vy[b] -= mb * dy; vz[b] -= mb * dz;
                    dx /= half_dt; dy /= half_dt; dz /= half_dt;
                    w[0] -= px * dx; w[1] -= px * dy; w[2] -= px * dz;
                    w[3] -= py * dx; w[4] -= py * dy; w[5] -= py * dz;
                    w[6] -= pz * dx; w[7] -= pz * dy; w[8] -= pz * dz;
                    done = 0;
                }
            }
            if (++iter > 1000) break;
        }
        if (!done) status[1] = 1;
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        double s = wave_sum(w[c]);
        if (lane == 0) s_red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double s = 0.0;
        if (threadIdx.x < 9)
            for (int q = 0; q < kRattleThreads / 64; ++q) s += s_red[q][threadIdx.x];
        partial[blockIdx.x * 16 + threadIdx.x] = s;
    }
}

// Momentum / kinetic-energy partial sums per block of 256 atoms:
// [sum m vx, m vy, m vz, m, m vx^2, m vy^2, m vz^2] over atoms with m > 0
// (kinetic_energy thermostat.c:34-63, remove_comm thermostat.c:69-108)
__global__ void __launch_bounds__(256)
k_mom_partial(int N, const double* __restrict__ mass, const double* __restrict__ v,
              double* __restrict__ partial) {
    __shared__ double s_red[4][7];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < N) {
        double m = mass[i];
        if (m > 0.0) {
            double vx = v[i], vy = v[N + i], vz = v[2 * N + i];
            a[0] = vx * m; a[1] = vy * m; a[2] = vz * m; a[3] = m;
            a[4] = m * vx * vx; a[5] = m * vy * vy; a[6] = m * vz * vz;
        }
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        double sm = wave_sum(a[c]);
        if (lane == 0) s_red[wave][c] = sm;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double sm = 0.0;
        if (threadIdx.x < 7)
            for (int q = 0; q < 4; ++q) sm += s_red[q][threadIdx.x];
        partial[blockIdx.x * 16 + threadIdx.x] = sm;
    }
}

// Taylor polynomial of exp of degree 4 for |x| < 2^-10: the first omitted
// term, x^5/120, stays below 7.4e-18 (a fifteenth of an ulp of 1).  Three
// dependent fused operations after x.
__device__ __forceinline__ double exp_small(double x) {
    const double x2 = x * x;
    const double a = fma(x, 1.0 / 6.0, 0.5);
    const double inner = fma(x2, 1.0 / 24.0, a);
    return fma(x2, inner, 1.0 + x);
}

// The chain's sweeps (thermostat.c:135-186).  SMALL uses exp_small for every
// exponential and returns false as soon as an argument leaves |x| < 2^-10
// (the caller then starts over with the library exp); along a trajectory the
// arguments, wdt/40 times a thermostat velocity, are of order 1e-4.  The
// chain is one thread's serial dependency, so every operation on it is paid
// at full latency.  The updates are written to keep it short: each step
// fraction (w dt / 8 Nc and the like) is formed once, v aa^2 + c aQ aa is
// fma(v aa, aa, (c aQ) aa), the forward pass's exponentials (of velocities
// the backward pass already fixed) leave the chain, and qMass v^2 - kT is one
// fma.  The results differ from the reference's operations by rounding only.
template <bool SMALL>
__device__ __forceinline__ bool nhc_sweeps(double vQ[CAPAC_NHC], double aQ[CAPAC_NHC], double dt, double kT,
                                           double NkT, double qMass, double inv_q, double eKsum,
                                           double& scale_out) {
    constexpr int Nc = 5, Nys = 5, M = CAPAC_NHC;
    const double w1 = 0.41449077179437571194, w3 = -0.65796308717750284778;
    // the step fractions of each weight
    const double c8[2] = {-w1 * dt / (8 * Nc), -w3 * dt / (8 * Nc)};
    const double c4[2] = {w1 * dt / (4 * Nc), w3 * dt / (4 * Nc)};
    const double c2[2] = {-w1 * dt / (2 * Nc), -w3 * dt / (2 * Nc)};
    double aa, x, scale = 1.0;
    bool ok = true;
#define NHC_EXP(arg) (x = (arg), SMALL ? (ok = ok && fabs(x) < 0x1p-10, exp_small(x)) : exp(x))
    for (int ic = 0; ic < Nc; ++ic) {
#pragma unroll
        for (int iys = 0; iys < Nys; ++iys) {
            const int w = iys == 2 ? 1 : 0;
            vQ[M - 1] = fma(aQ[M - 1], c4[w], vQ[M - 1]);
#pragma unroll
            for (int im = 0; im < M - 1; ++im) {
                const int k = M - 1 - (im + 1);
                aa = NHC_EXP(c8[w] * vQ[k + 1]);
                vQ[k] = fma(vQ[k] * aa, aa, (c4[w] * aQ[k]) * aa);
            }
            aa = NHC_EXP(c2[w] * vQ[0]);
            scale *= aa;
            aQ[0] = fma(eKsum * scale, scale, -NkT) * inv_q;
            double ef[M - 1];
#pragma unroll
            for (int im = 0; im < M - 1; ++im) ef[im] = NHC_EXP(c8[w] * vQ[im + 1]);
#pragma unroll
            for (int im = 0; im < M - 1; ++im) {
                vQ[im] = fma(vQ[im] * ef[im], ef[im], (c4[w] * ef[im]) * aQ[im]);
                aQ[im + 1] = fma(qMass * vQ[im], vQ[im], -kT) * inv_q;
            }
            vQ[M - 1] = fma(aQ[M - 1], c4[w], vQ[M - 1]);
        }
        if (SMALL && !ok) return false;
    }
#undef NHC_EXP
    scale_out = scale;
    return true;
}

// The centre-of-mass velocity and the kinetic energy after the velocities
// are scaled by s and the centre-of-mass motion is removed (remove_comm +
// kinetic_energy, thermostat.c:34-108), from the sums of the unscaled
// velocities: with P = sum m v and S = sum m v^2 per component over the atoms
// with mass M, the removal subtracts c = s P / M and leaves
// sum m (s v - c)^2 = s^2 (S - P^2 / M) per component.  One pass over the
// atoms (k_scale_com) then applies s and c, where the reference makes three
// (scale, momentum sum and removal, kinetic sum).  The values differ from
// the reference's sums by rounding only: P is tiny once the motion has been
// removed at the previous step, so the difference S - P^2/M cancels nothing.
// md[26..28] = c, md[0] = eKsum, md[1] = temperature.
__device__ __forceinline__ void com_kinetic_from_sums(double* __restrict__ md, const double* __restrict__ sums,
                                                      double scale, int ndf) {
    const double msum = sums[3];
    double ek = 0.0;
    for (int c = 0; c < 3; ++c) {
        const double pc = sums[c];
        md[26 + c] = scale * pc / msum;
        ek += scale * scale * (sums[4 + c] - pc * pc / msum);
    }
    md[0] = ek;
    md[1] = ek / (K_BOLTZ * ndf);
}

// nose_hoover_chain (thermostat.c:117-191); the velocity scale goes to md[22].
// The chain lives in registers (all indices are compile-time after unrolling).
// With sums (the step's second half), eKsum comes from the momentum sums and
// the kinetic energy after the scale and the centre-of-mass removal is
// formed as com_kinetic_from_sums describes.  The divisions by qMass are
// products with its reciprocal, and small exponentials use exp_small: both
// differ from the reference's operations by rounding only.
__device__ __forceinline__ void nhc_body(double* __restrict__ md, double dt, double kT, int ndf,
                                         const double* __restrict__ sums) {
    constexpr int M = CAPAC_NHC;
    if (sums) md[0] = sums[4] + sums[5] + sums[6];
    double vQ[M], aQ[M];
#pragma unroll
    for (int k = 0; k < M; ++k) vQ[k] = md[2 + k];
    const double qMass = md[25], eKsum = md[0];
    const double inv_q = 1.0 / qMass;
    const double NkT = kT * ndf;
    aQ[0] = (eKsum - NkT) * inv_q;
#pragma unroll
    for (int im = 0; im < M - 1; ++im) aQ[im + 1] = (qMass * vQ[im] * vQ[im] - kT) * inv_q;
    double vQ0[M], aQ0[M];
#pragma unroll
    for (int k = 0; k < M; ++k) { vQ0[k] = vQ[k]; aQ0[k] = aQ[k]; }
    double scale = 1.0;
    if (!nhc_sweeps<true>(vQ, aQ, dt, kT, NkT, qMass, inv_q, eKsum, scale)) {
#pragma unroll
        for (int k = 0; k < M; ++k) { vQ[k] = vQ0[k]; aQ[k] = aQ0[k]; }
        nhc_sweeps<false>(vQ, aQ, dt, kT, NkT, qMass, inv_q, eKsum, scale);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) { md[2 + k] = vQ[k]; md[12 + k] = aQ[k]; }
    md[22] = scale;
    if (sums) {
        com_kinetic_from_sums(md, sums, scale, ndf);
    } else {
        md[0] = eKsum * (scale * scale);
        md[1] *= scale * scale;
    }
}

__global__ void k_nhc(double* __restrict__ md, double dt, double kT, int ndf, const double* __restrict__ sums) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    nhc_body(md, dt, kT, ndf, sums);
}

// The step's second-half chain with the momentum sums' reduction in the same
// launch (when no other reduction is queued): the workgroup forms red[96..103)
// exactly as k_reduce does, then one thread runs the chain on them.
__global__ void __launch_bounds__(kRedThreads)
k_nhc_sums(double* __restrict__ md, double dt, double kT, int ndf, const double* __restrict__ partial, int nblk,
           double* __restrict__ sums_out) {
    __shared__ double s_w[16][16];
    __shared__ double s_sums[16];
    reduce_rows_block(partial, nblk, 7, sums_out, s_w, s_sums);
    __syncthreads();
    if (threadIdx.x == 0) nhc_body(md, dt, kT, ndf, s_sums);
}

// The same, then the next step's first-half chain at kT_next on the same
// thread (it needs only the chain and the kinetic energy this chain leaves):
// this step's scale moves to md[29] for the next step's first pass, which
// also applies this step's centre-of-mass removal (k_scale_kick_drift with
// pre_com); md[22] holds the next step's scale.
__global__ void __launch_bounds__(kRedThreads)
k_nhc_sums_next(double* __restrict__ md, double dt, double kT, double kT_next, int ndf,
                const double* __restrict__ partial, int nblk, double* __restrict__ sums_out) {
    __shared__ double s_w[16][16];
    __shared__ double s_sums[16];
    reduce_rows_block(partial, nblk, 7, sums_out, s_w, s_sums);
    __syncthreads();
    if (threadIdx.x == 0) {
        nhc_body(md, dt, kT, ndf, s_sums);
        md[29] = md[22];
        nhc_body(md, dt, kT_next, ndf, NULL);
    }
}

// the same without a thermostat (scale 1)
__global__ void k_com_finish(double* __restrict__ md, const double* __restrict__ sums, int ndf) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    com_kinetic_from_sums(md, sums, 1.0, ndf);
}

// v = s v - c for atoms with mass, v = s v for virtual sites (the
// thermostat scales every atom, the removal skips massless ones); s = md[22]
// when scale is set, else 1; c = md[26..28]
__global__ void __launch_bounds__(256)
k_scale_com(int N, const double* __restrict__ mass, double* __restrict__ v, const double* __restrict__ md,
            int scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double s = scale ? md[22] : 1.0;
    const bool has_mass = mass[i] > 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t k = (size_t)c * N + i;
        double vk = scale ? v[k] * s : v[k];
        if (has_mass) vk -= md[26 + c];
        v[k] = vk;
    }
}

// apply_pbc (force.c:127-165): wrap the middle atom of each molecule into the
// box, the others to within half a box of it
__global__ void k_wrap(int n_mols, int N, BoxConst bc, const int* __restrict__ first,
                       const int* __restrict__ last, double* __restrict__ r) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_mols) return;
    int a0 = first[m], a1 = last[m], mid = (a0 + a1) / 2;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double* x = r + c * N;
        double L = bc.box[c], h = 0.5 * bc.box[c];
        if (x[mid] > L) x[mid] -= L; else if (x[mid] < 0.0) x[mid] += L;
        for (int i = a0; i <= a1; ++i) {
            if (i == mid) continue;
            if (x[i] - x[mid] > h) x[i] -= L; else if (x[i] - x[mid] < -h) x[i] += L;
        }
    }
}

// get_fmax_rms (force.c:361-378)
__global__ void __launch_bounds__(kRedThreads)
k_fmax(int N, const double* __restrict__ f, double* __restrict__ md) {
    __shared__ double s_w[16];
    __shared__ double s_m[16];
    double mx = 0.0, sum = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double f2 = f[i] * f[i] + f[N + i] * f[N + i] + f[2 * N + i] * f[2 * N + i];
        sum += f2;
        mx = fmax(mx, f2);
    }
    sum = block_sum1024(sum, s_w);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, s_m[w]);
        md[23] = sqrt(m);
        md[24] = sqrt(sum / N);
    }
}

}  // namespace

static inline int grid(int n, int b) { return (n + b - 1) / b; }

int md_kick(capac_ctx* c) {
    k_kick<<<grid(c->bc.N, 256), 256, 0, c->stream>>>(c->bc.N, c->bc.Nm, c->fix_metal, 0.5 * c->dt,
                                                      c->f, c->inv_mass, c->v);
    CHECK_HIP(hipGetLastError());
    return 0;
}

int md_rattle1(capac_ctx* c) {
    if (c->n_cstr == 0) return 0;
    k_rattle1<<<grid(c->bc.n_mols, kRattleThreads), kRattleThreads, 0, c->stream>>>(c->bc.n_mols, c->bc.N, c->bc, c->dt,
                                                              c->mol_cstr, c->mol_first, c->mol_last,
                                                              c->cstr_i, c->cstr_j,
                                                              c->cstr_r, c->inv_mass, c->r, c->v,
                                                              c->old_r, c->md_status);
    CHECK_HIP(hipGetLastError());
    return 0;
}

// red[64..73) = virial of the velocity constraints; with_kick: the half
// kick first, in the same launch (with constraints; else its own launch)
int md_rattle2(capac_ctx* c, bool with_kick) {
    if (c->n_cstr == 0) {
        if (with_kick) {
            int rc = md_kick(c);
            if (rc) return rc;
        }
        return fill_zero_d(c, c->red + 64, 16);
    }
    int nblk = grid(c->bc.n_mols, kRattleThreads);
    k_rattle2<<<nblk, kRattleThreads, 0, c->stream>>>(c->bc.n_mols, c->bc.N, c->bc, c->dt, c->mol_cstr,
                                           c->mol_first, c->mol_last, c->cstr_i, c->cstr_j, c->cstr_r, c->inv_mass, c->r, c->v,
                                           c->md_status, partial_seg(c, SEG_RATTLE), c->bc.Nm, c->fix_metal,
                                           with_kick ? c->f : NULL);
    CHECK_HIP(hipGetLastError());
    return reduce_energies(c, SEG_RATTLE, nblk, 9, 64);
}

// red[96..103) = momentum / kinetic sums of the current velocities
static int mom_sums(capac_ctx* c) {
    int nblk = grid(c->bc.N, 256);
    k_mom_partial<<<nblk, 256, 0, c->stream>>>(c->bc.N, c->mass, c->v, partial_seg(c, SEG_MOM));
    CHECK_HIP(hipGetLastError());
    int rc = reduce_partials(c, SEG_MOM, nblk, 7, 96);
    return rc ? rc : flush_reductions(c);      // read on the device right after
}

// applies the scale (md[22], when `scale`) and the centre-of-mass removal
// that k_nhc or k_com_finish formed from the sums
static int scale_com(capac_ctx* c, int scale) {
    k_scale_com<<<grid(c->bc.N, 256), 256, 0, c->stream>>>(c->bc.N, c->mass, c->v, c->md, scale);
    CHECK_HIP(hipGetLastError());
    return 0;
}

// remove_comm and kinetic_energy (mpi-main.c:632-633 after the first force
// call; the end of a step without thermostat)
int md_remove_com_kinetic(capac_ctx* c) {
    int rc = mom_sums(c);
    if (rc) return rc;
    k_com_finish<<<1, 64, 0, c->stream>>>(c->md, c->red + 96, c->ndf);
    CHECK_HIP(hipGetLastError());
    return scale_com(c, 0);
}

// First half of a step (mpi-main.c:885-925): nvt chain, scale, half kick
// and drift, in two launches (k_nhc, k_scale_kick_drift); one when the
// previous step's last launch already ran this chain (c->com_pending: its
// centre-of-mass pass then runs here first)
int md_first_half(capac_ctx* c, double kT) {
    const int pre = c->com_pending;
    c->com_pending = 0;
    if (c->nvt && !pre) {
        k_nhc<<<1, 64, 0, c->stream>>>(c->md, c->dt, kT, c->ndf, NULL);
        CHECK_HIP(hipGetLastError());
    }
    k_scale_kick_drift<<<grid(c->bc.N, 256), 256, 0, c->stream>>>(c->bc.N, c->bc.Nm, c->fix_metal, c->nvt,
                                                                  0.5 * c->dt, c->dt, c->md, c->f,
                                                                  c->inv_mass, c->v, c->r, c->old_r, pre,
                                                                  c->mass);
    CHECK_HIP(hipGetLastError());
    return 0;
}

// Second half after RATTLE_2 (mpi-main.c:955-973): kinetic energy, nvt
// chain and scale, centre-of-mass removal, kinetic energy, in three
// launches after the sums: one momentum-sum pass, the chain (which also forms
// the centre-of-mass velocity and the final kinetic energy from the sums)
// and one pass applying the scale and the removal.
int md_second_half(capac_ctx* c, double kT, double kT_next) {
    if (!c->nvt) return md_remove_com_kinetic(c);
    if (c->n_pending == 0) {
        // nothing else queued (a step that forms no energies): the chain's
        // launch reduces the sums itself.  kT_next > 0 (an unobserved step
        // with a next step in the same call): that launch also runs the next
        // step's first-half chain, and the next step's first pass applies
        // this step's scale and centre-of-mass removal.
        int nblk = grid(c->bc.N, 256);
        if (nblk > c->seg_cap[SEG_MOM]) {
            capac_set_error("internal: reduction of %d blocks exceeds segment %d", nblk, SEG_MOM);
            return CAPAC_ERR_HIP;
        }
        k_mom_partial<<<nblk, 256, 0, c->stream>>>(c->bc.N, c->mass, c->v, partial_seg(c, SEG_MOM));
        if (kT_next > 0.0) {
            k_nhc_sums_next<<<1, kRedThreads, 0, c->stream>>>(c->md, c->dt, kT, kT_next, c->ndf,
                                                              partial_seg(c, SEG_MOM), nblk, c->red + 96);
            CHECK_HIP(hipGetLastError());
            c->com_pending = 1;
            return 0;
        }
        k_nhc_sums<<<1, kRedThreads, 0, c->stream>>>(c->md, c->dt, kT, c->ndf, partial_seg(c, SEG_MOM), nblk,
                                                     c->red + 96);
        CHECK_HIP(hipGetLastError());
        return scale_com(c, 1);
    }
    int rc = mom_sums(c);
    if (rc) return rc;
    k_nhc<<<1, 64, 0, c->stream>>>(c->md, c->dt, kT, c->ndf, c->red + 96);
    CHECK_HIP(hipGetLastError());
    return scale_com(c, 1);
}

int md_wrap(capac_ctx* c) {
    k_wrap<<<grid(c->bc.n_mols, 256), 256, 0, c->stream>>>(c->bc.n_mols, c->bc.N, c->bc, c->mol_first,
                                                           c->mol_last, c->r);
    CHECK_HIP(hipGetLastError());
    return 0;
}

int md_fmax(capac_ctx* c) {
    k_fmax<<<1, kRedThreads, 0, c->stream>>>(c->bc.N, c->f, c->md);
    CHECK_HIP(hipGetLastError());
    return 0;
}

