/* cpu_ref.c -- CPU oracle: FP64 restatement of the reference's force + CPIM
 * path and MD step (TEST INFRASTRUCTURE ONLY; never linked by the product).
 *
 * Every function keeps the reference's loop order, summation order and
 * operand association so that, built with gcc -O2 -ffp-contract=off on
 * x86-64, it reproduces the reference bit for bit.  Constants are the
 * reference's own literals (typedef.h:30-93, cpff.c:72-75,758-759), including
 * its truncated 1/sqrt(pi).
 *
 * Parity pins (tests/test_oracle.py):
 *   - golden vectors dumped from the reference compiled from its own sources
 *     (oracle/ref_dump.cpp over tests/golden decks): nonbonded, Sutton-Chen,
 *     CPIM right-hand side, COO relay matrix, BiCGSTAB solution, CPIM forces,
 *     the full force call and three MD steps -- bitwise;
 *   - mathematics: DSF/LJ force and energy vanish at the cut-off, central
 *     finite differences of every energy term (forces = -gradient, with the
 *     CPIM energy re-solved tightly), relay-matrix symmetry and its r -> inf
 *     limit, the solve against a dense LU solve, charge neutrality per
 *     nanoparticle, Newton's third law of the pair terms.
 * Built with -DORACLE_OMP it parallelises the reference's own decomposition
 * units (i atoms, metal rows) with OpenMP for the all-core CPU baseline;
 * that build changes summation order and is used for timing only.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef ORACLE_OMP
#include <omp.h>
#endif

/* reference literals (typedef.h:44,67,73,77,85,89-93; cpff.c:72,75,758,759) */
#define R_INV_SQRT_PI 0.564189583547756
#define R_FQQ 138.935457696552
#define R_NM2BOHR 18.8972612456506
#define R_K_BOLTZ 0.00831446214546895
#define R_MAX_ITER 1000
#define R_THRSHD 1.0e-15
#define R_CONV_E 2.01552955019358e-05
#define R_CONV_V 3.80879884583368e-04
#define R_FQQ_AU2 49614.7526045429
#define R_FQQ_AU1 2625.49964037578
#define R_PI 3.14159265358979323846

struct oref {
    capac_params p;
    int N, Nm, n_mat, n_np;
    double box[6];
    double rc, rc2, alpha, a2_sqrtPI, wolfConst, erfc_arCut, inv_rc12, inv_rc6;
    double inv_polar, inv_capac, inv_R_pp, inv_R_pq, inv_R_qq;
    /* COO relay matrix in reference order; group_start[i] is the first
       entry of metal row-group i (its rows 3i..3i+2, 3Nm+i and its Lagrange
       entries), group_start[Nm] the first of the nanoparticle rows */
    long nnz, cap;
    long *row, *col;
    double *val, *diag;
    long* group_start;
    double *vec_ext, *vec_pq, *isd;
    double *Ax, *r0, *r, *pp, *v, *s, *t, *y, *z, *Kt, *K;
    /* state, 3N SoA */
    double *rr, *vv, *ff, *old_r;
    double pot[15], vir[9];
    /* thermostat */
    double vQ[CAPAC_NHC], aQ[CAPAC_NHC];
    double eKsum, inst_temp, first_temp, ext_temp, ref_temp, kT, qMass;
    double fmax, frms;
    int solve_iters, solve_kind;
    double solve_rTr, solve_bTb;
    int int_pair_count;
    double timers[15];
    double solve_crit;          /* stop test rTr < solve_crit * bTb */
    double *em_s, *em_old_f;    /* minimiser direction and previous forces */
};

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void* zalloc(size_t n, size_t sz) {
    void* q = calloc(n ? n : 1, sz);
    if (!q) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
    return q;
}

/* pbc_12 (force.c:111-121): single shift against the half box */
static inline void pbc12(const double* box, double* x, double* y, double* z) {
    if (*x < -box[3]) *x += box[0]; else if (*x > box[3]) *x -= box[0];
    if (*y < -box[4]) *y += box[1]; else if (*y > box[4]) *y -= box[1];
    if (*z < -box[5]) *z += box[2]; else if (*z > box[5]) *z -= box[2];
}

/* dist_2 (force.c:41-44) */
static inline double dist2(double x, double y, double z) { return x * x + y * y + z * z; }

/* pbc (force.c:104-109) */
static inline double pbc1(double dx, double L) {
    if (dx < -0.5 * L) dx += L; else if (dx > 0.5 * L) dx -= L;
    return dx;
}

oref* oref_create(const capac_params* p) {
    oref* o = zalloc(1, sizeof(oref));
    o->p = *p;
    o->solve_crit = 1.0e-08;
    int N = o->N = p->n_atoms;
    o->Nm = p->n_metal;
    o->n_np = p->n_nps;
    o->n_mat = p->n_metal > 0 ? 4 * p->n_metal + p->n_nps : 0;
    for (int a = 0; a < 3; ++a) { o->box[a] = p->box[a]; o->box[3 + a] = p->box[a] * 0.5; }
    /* derived constants (mpi-main.c:170-178, 202-203, 315-323) */
    o->rc = p->rcut;
    o->rc2 = p->rcut * p->rcut;
    o->alpha = p->dsf_alpha;
    o->a2_sqrtPI = o->alpha * 2.0 * R_INV_SQRT_PI;
    o->wolfConst = erfc(o->alpha * o->rc) / o->rc2 +
                   o->a2_sqrtPI * exp(-o->alpha * o->alpha * o->rc2) / o->rc;
    o->erfc_arCut = erfc(o->alpha * o->rc) / o->rc;
    o->inv_rc12 = 1.0 / pow(o->rc, 12.0);
    o->inv_rc6 = 1.0 / pow(o->rc, 6.0);
    o->inv_polar = 1.0 / p->cpff_polar;
    o->inv_capac = 1.0 / p->cpff_capac;
    double R_q = sqrt(2.0 / R_PI) * p->cpff_capac;
    double R_p = pow(sqrt(2.0 / R_PI) * p->cpff_polar / 3.0, 1.0 / 3.0);
    o->inv_R_qq = 1.0 / sqrt(R_q * R_q + R_q * R_q);
    o->inv_R_pq = 1.0 / sqrt(R_p * R_p + R_q * R_q);
    o->inv_R_pp = 1.0 / sqrt(R_p * R_p + R_p * R_p);

    int nm = o->n_mat > 0 ? o->n_mat : 1;
    o->vec_ext = zalloc(nm, sizeof(double));
    o->vec_pq = zalloc(nm, sizeof(double));
    o->diag = zalloc(nm, sizeof(double));
    for (int i = 0; i < nm; ++i) o->diag[i] = 1.0;
    o->isd = zalloc(o->Nm + 1, sizeof(double));
    double** vecs[11] = {&o->Ax, &o->r0, &o->r, &o->pp, &o->v, &o->s, &o->t, &o->y, &o->z, &o->Kt, &o->K};
    for (int i = 0; i < 11; ++i) *vecs[i] = zalloc(nm, sizeof(double));
    o->rr = zalloc(3 * (size_t)N, sizeof(double));
    o->vv = zalloc(3 * (size_t)N, sizeof(double));
    o->ff = zalloc(3 * (size_t)N, sizeof(double));
    o->old_r = zalloc(3 * (size_t)N, sizeof(double));
    o->ref_temp = p->ref_temp;
    o->ext_temp = p->ref_temp;
    o->kT = R_K_BOLTZ * p->ref_temp;
    o->qMass = (double)p->ndf * o->kT * p->tau_t * p->tau_t;   /* mpi-main.c:551-552 */
    return o;
}

void oref_destroy(oref* o) {
    if (!o) return;
    free(o->row); free(o->col); free(o->val); free(o->diag); free(o->group_start);
    free(o->vec_ext); free(o->vec_pq); free(o->isd);
    double* vecs[11] = {o->Ax, o->r0, o->r, o->pp, o->v, o->s, o->t, o->y, o->z, o->Kt, o->K};
    for (int i = 0; i < 11; ++i) free(vecs[i]);
    free(o->rr); free(o->vv); free(o->ff); free(o->old_r);
    free(o->em_s); free(o->em_old_f);
    free(o);
}

int oref_n_mat(const oref* o) { return o->n_mat; }

void oref_set_state(oref* o, const double* r, const double* v, const double* x) {
    size_t n3 = 3 * (size_t)o->N;
    if (r) memcpy(o->rr, r, n3 * sizeof(double));
    if (v) memcpy(o->vv, v, n3 * sizeof(double));
    if (x && o->n_mat) memcpy(o->vec_pq, x, o->n_mat * sizeof(double));
}

void oref_get_state(const oref* o, double* r, double* v, double* x, double* f) {
    size_t n3 = 3 * (size_t)o->N;
    if (r) memcpy(r, o->rr, n3 * sizeof(double));
    if (v) memcpy(v, o->vv, n3 * sizeof(double));
    if (f) memcpy(f, o->ff, n3 * sizeof(double));
    if (x && o->n_mat) memcpy(x, o->vec_pq, o->n_mat * sizeof(double));
}

void oref_set_int_pair_count(oref* o, int on) { o->int_pair_count = on; }

/* Fitting parameters (alpha, c, R_q, R_p, eps_SC, c_SC) with the widths
 * free instead of derived from alpha and c (mpi-main.c:318-323); the
 * checker for the GPU's gradients evaluates its loss through this. */
void oref_set_theta(oref* o, const double* t) {
    o->inv_polar = 1.0 / t[0];
    o->inv_capac = 1.0 / t[1];
    o->inv_R_qq = 1.0 / sqrt(t[2] * t[2] + t[2] * t[2]);
    o->inv_R_pq = 1.0 / sqrt(t[3] * t[3] + t[2] * t[2]);
    o->inv_R_pp = 1.0 / sqrt(t[3] * t[3] + t[3] * t[3]);
    o->p.qsc_eps = t[4];
    o->p.qsc_c = t[5];
}

/* last pair index of rank 0 of 1 with the reference's int product
 * nAtoms*(nAtoms-1)/2 (mpi-main.c:406), which wraps above 46,341 atoms */
long oref_reference_pair_end(int n) {
    unsigned int prod = (unsigned int)n * (unsigned int)(n - 1);
    int wrapped = (int)prod;
    long n_pairs = wrapped / 2;
    return n_pairs - 1;
}

/* ------------------------------------------------------------------ pairs */

/* compute_nonbonded (force.c:710-816) for funct 1, non-1-4 pair; returns the
 * force factors of the vdW and Coulomb parts (fj = fij * r) */
static inline void pair_lj_dsf(const oref* o, double rij2, double c6, double c12, double qi,
                               double qj, double* flj, double* fqq, double* pot) {
    double rij = sqrt(rij2);
    *flj = 0.0;
    *fqq = 0.0;
    if (c6 != 0.0 || c12 != 0.0) {
        double rij6 = rij2 * rij2 * rij2;
        double rij12 = rij6 * rij6;
        if (o->p.use_vdw == 1) {
            double f_r = 12.0 * c12 / rij12 - 6.0 * c6 / rij6;
            double f_c = 12.0 * c12 * o->inv_rc12 - 6.0 * c6 * o->inv_rc6;
            *flj = (f_r - f_c) / rij2;
            pot[8] += ((c12 / rij12 - c6 / rij6) + (rij - o->rc) * f_c / o->rc -
                       (c12 * o->inv_rc12 - c6 * o->inv_rc6));
        } else {
            *flj = (c12 * 2.0 / rij12 - c6 / rij6) * 6.0 / rij2 * 1.0;
            pot[8] += (c12 / rij12 - c6 / rij6) * 1.0;
        }
    }
    if (qi != 0.0 && qj != 0.0) {
        if (o->p.use_coulomb == 1) {
            double erfc_arij = erfc(o->alpha * rij);
            *fqq = R_FQQ * qi * qj *
                   ((erfc_arij / rij2 + o->a2_sqrtPI * exp(-o->alpha * o->alpha * rij2) / rij) -
                    o->wolfConst) / rij;
            pot[7] += R_FQQ * qi * qj * (erfc_arij / rij - o->erfc_arCut + o->wolfConst * (rij - o->rc));
        } else {
            *fqq = 1.0 * R_FQQ * qi * qj / (rij2 * rij);
            pot[7] += 1.0 * R_FQQ * qi * qj / rij;
        }
    }
}

/* compute_erf_vdw (force.c:591-632) */
static inline double pair_erf_vdw(double rij2, double c6, double c12, double width, double* pot) {
    double rij = sqrt(rij2);
    double fij = 0.0;
    if (c6 != 0.0 || c12 != 0.0) {
        double rij6 = rij2 * rij2 * rij2;
        double rij12 = rij6 * rij6;
        double cr12 = c12 / rij12;
        double cr6 = c6 / rij6;
        double rR = rij / width;
        double erf_rR = erf(rR);
        double vij = (cr12 - cr6) * erf_rR;
        fij = (12.0 * cr12 - 6.0 * cr6) / rij * erf_rR -
              (cr12 - cr6) * 2.0 * R_INV_SQRT_PI * exp(-rR * rR) / width;
        fij /= rij;
        pot[8] += vij;
    }
    return fij;
}

/* compute_morse (force.c:639-668) */
static inline double pair_morse(double rij2, double D, double a, double R, double* pot) {
    double rij = sqrt(rij2);
    double e = exp(-a * (rij - R));
    double e2 = e * e;
    double fij = 2.0 * D * a * (e2 - e) / rij;
    pot[8] += D * (e2 - 2.0 * e);
    return fij;
}

/* compute_buckingham (force.c:674-703) */
static inline double pair_buckingham(double rij2, double A, double B, double c6, double* pot) {
    double rij = sqrt(rij2);
    double rij6 = rij2 * rij2 * rij2;
    double rij8 = rij6 * rij2;
    double e = exp(-B * rij);
    double fij = A * e * B / rij - 6.0 * c6 / rij8;
    pot[8] += (A * e - c6 / rij6);
    return fij;
}

/* one accepted pair of mpi_nonb_pair, forces and virial (force.c:1476-1537) */
static inline void nonb_one(const oref* o, int i, int j, double rx, double ry, double rz,
                            double rij2, double* f, double* pot, double* vir) {
    const capac_params* p = &o->p;
    int N = o->N, nt = p->n_types;
    int ij = p->atomtype[i] * nt + p->atomtype[j];
    int funct = p->nb_funct[ij];
    double flj = 0.0, fqq = 0.0;
    if (funct == 4) {
        flj = pair_erf_vdw(rij2, p->nb_c6[ij], p->nb_c12[ij], p->nb_a[ij], pot);
    } else if (funct == 3) {
        flj = pair_morse(rij2, p->nb_a[ij], p->nb_b[ij], p->nb_c6[ij], pot);
    } else if (funct == 2) {
        flj = pair_buckingham(rij2, p->nb_a[ij], p->nb_b[ij], p->nb_c6[ij], pot);
    } else if (funct == 1) {
        pair_lj_dsf(o, rij2, p->nb_c6[ij], p->nb_c12[ij], p->charge[i], p->charge[j], &flj, &fqq, pot);
    }
    double ljx = flj * rx, ljy = flj * ry, ljz = flj * rz;
    double qqx = fqq * rx, qqy = fqq * ry, qqz = fqq * rz;
    /* virial_vec_nb (force.c:562-583) */
    double fx = ljx + qqx, fy = ljy + qqy, fz = ljz + qqz;
    vir[0] -= rx * fx; vir[1] -= rx * fy; vir[2] -= rx * fz;
    vir[3] -= ry * fx; vir[4] -= ry * fy; vir[5] -= ry * fz;
    vir[6] -= rz * fx; vir[7] -= rz * fy; vir[8] -= rz * fz;
    f[i] += -ljx + -qqx;
    f[N + i] += -ljy + -qqy;
    f[2 * N + i] += -ljz + -qqz;
    f[j] += ljx + qqx;
    f[N + j] += ljy + qqy;
    f[2 * N + j] += ljz + qqz;
}

/* mpi_nonb_pair (force.c:1402-1540) at one rank: all pairs j < i in order */
void oref_nonb(oref* o, const double* r, double* f, double* pot, double* vir) {
    const capac_params* p = &o->p;
    int N = o->N;
    const double *rx = r, *ry = r + N, *rz = r + 2 * N;
    long pair_end = o->int_pair_count ? oref_reference_pair_end(N) : (long)N * (N - 1) / 2 - 1;
    double t0 = now();
#ifdef ORACLE_OMP
    (void)pair_end;
    int nth = omp_get_max_threads();
    double* fb = zalloc((size_t)nth * 3 * N, sizeof(double));
    double* pb = zalloc((size_t)nth * 24, sizeof(double));
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
        double* fl = fb + (size_t)tid * 3 * N;
        double* pl = pb + (size_t)tid * 24;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < i; ++j) {
                if (p->mol_id[j] == p->mol_id[i] && ((p->excl_mask[i] >> p->local_idx[j]) & 1u)) continue;
                double x = rx[j] - rx[i], y = ry[j] - ry[i], z = rz[j] - rz[i];
                pbc12(o->box, &x, &y, &z);
                if (x >= o->rc && y >= o->rc && z >= o->rc) continue;
                double r2 = dist2(x, y, z);
                if (r2 >= o->rc2) continue;
                nonb_one(o, i, j, x, y, z, r2, fl, pl, pl + 15);
            }
    }
    for (int t = 0; t < nth; ++t) {
        for (size_t k = 0; k < 3 * (size_t)N; ++k) f[k] += fb[(size_t)t * 3 * N + k];
        pot[7] += pb[t * 24 + 7];
        pot[8] += pb[t * 24 + 8];
        for (int k = 0; k < 9; ++k) vir[k] += pb[t * 24 + 15 + k];
    }
    free(fb); free(pb);
#else
    long count = 0;
    for (int i = 0; i < N; ++i) {
        if (count > pair_end) break;
        for (int j = 0; j < i; ++j) {
            if (count > pair_end) break;
            ++count;
            if (p->mol_id[j] == p->mol_id[i] && ((p->excl_mask[i] >> p->local_idx[j]) & 1u)) continue;
            double x = rx[j] - rx[i], y = ry[j] - ry[i], z = rz[j] - rz[i];
            pbc12(o->box, &x, &y, &z);
            if (x >= o->rc && y >= o->rc && z >= o->rc) continue;
            double r2 = dist2(x, y, z);
            if (r2 >= o->rc2) continue;
            nonb_one(o, i, j, x, y, z, r2, f, pot, vir);
        }
    }
#endif
    o->timers[4] += now() - t0;
}

/* ------------------------------------------------------------------ bonded */

typedef struct { double x, y, z; } ovec;

static inline ovec ov_pbc(const oref* o, double dx, double dy, double dz) {
    ovec v = {pbc1(dx, o->box[0]), pbc1(dy, o->box[1]), pbc1(dz, o->box[2])};
    return v;
}
static inline double ov_dot(ovec a, ovec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline ovec ov_scale(double s, ovec v) { ovec r = {s * v.x, s * v.y, s * v.z}; return r; }
static inline ovec ov_cross(ovec a, ovec b) {
    ovec r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
/* virial_rvec_fvec_full (force.c:84-100) */
static inline void ov_virial(ovec r, ovec f, double* vir) {
    vir[0] -= r.x * f.x; vir[1] -= r.x * f.y; vir[2] -= r.x * f.z;
    vir[3] -= r.y * f.x; vir[4] -= r.y * f.y; vir[5] -= r.y * f.z;
    vir[6] -= r.z * f.x; vir[7] -= r.z * f.y; vir[8] -= r.z * f.z;
}
static inline void ov_add(double* f, int N, int i, ovec v) {
    f[i] += v.x; f[N + i] += v.y; f[2 * N + i] += v.z;
}

/* compute_bond_1 / compute_bond_3 (force.c:822-909): returns fj, fi = -fj */
static ovec bond_term(const oref* o, const double* r, int i, int j, int funct, const double* bp,
                      double* pot, double* vir) {
    int N = o->N;
    ovec d = ov_pbc(o, r[j] - r[i], r[N + j] - r[N + i], r[2 * N + j] - r[2 * N + i]);
    double rij = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    double fij;
    if (funct == 1) {
        double b0 = bp[0], kb = bp[1];
        fij = kb * (b0 - rij) / rij;
        pot[2] += 0.5 * kb * (b0 - rij) * (b0 - rij);
    } else {
        double b0 = bp[0], D = bp[1], beta = bp[2];
        double e = exp(-beta * (rij - b0));
        fij = -2.0 * D * beta * e * (1.0 - e) / rij;
        pot[2] += D * pow(1.0 - e, 2.0);
    }
    ovec fj = {fij * d.x, fij * d.y, fij * d.z};
    vir[0] += -d.x * fj.x; vir[1] += -d.x * fj.y; vir[2] += -d.x * fj.z;
    vir[3] += -d.y * fj.x; vir[4] += -d.y * fj.y; vir[5] += -d.y * fj.z;
    vir[6] += -d.z * fj.x; vir[7] += -d.z * fj.y; vir[8] += -d.z * fj.z;
    return fj;
}

/* mpi_bonded (force.c:1694-1904) over all molecules, one rank */
void oref_bonded(oref* o, const double* r, double* f, double* pot, double* vir) {
    const capac_params* p = &o->p;
    int N = o->N, M = p->n_mols, M1 = M + 1;
    if (!p->mol_terms) return;
    const int* rg = p->mol_terms;
    for (int m = 0; m < M; ++m) {
        /* bonds */
        for (int t = rg[m]; t < rg[m + 1]; ++t) {
            int i = p->bond_atoms[2 * t], j = p->bond_atoms[2 * t + 1];
            ovec fj = bond_term(o, r, i, j, p->bond_funct[t], p->bond_par + 3 * t, pot, vir);
            f[i] += -fj.x; f[N + i] += -fj.y; f[2 * N + i] += -fj.z;
            ov_add(f, N, j, fj);
        }
        /* 1-4 pairs: compute_nonbonded with is_14_pair = 1 (force.c:710-816) */
        for (int t = rg[M1 + m]; t < rg[M1 + m + 1]; ++t) {
            int i = p->pair14_atoms[2 * t], j = p->pair14_atoms[2 * t + 1];
            const double* pp = p->pair14_par + 6 * t;
            double c6 = pp[0], c12 = pp[1], qi = pp[2], qj = pp[3], sLJ = pp[4], sQQ = pp[5];
            double rx = pbc1(r[j] - r[i], o->box[0]);
            double ry = pbc1(r[N + j] - r[N + i], o->box[1]);
            double rz = pbc1(r[2 * N + j] - r[2 * N + i], o->box[2]);
            double rij2 = rx * rx + ry * ry + rz * rz;
            double rij = sqrt(rij2);
            ovec lj = {0.0, 0.0, 0.0}, qq = {0.0, 0.0, 0.0};
            if (c6 != 0.0 || c12 != 0.0) {
                double rij6 = rij2 * rij2 * rij2;
                double rij12 = rij6 * rij6;
                double fij = (c12 * 2.0 / rij12 - c6 / rij6) * 6.0 / rij2 * sLJ;
                pot[8] += (c12 / rij12 - c6 / rij6) * sLJ;
                lj.x = fij * rx; lj.y = fij * ry; lj.z = fij * rz;
            }
            if (qi != 0.0 && qj != 0.0) {
                double fij = sQQ * R_FQQ * qi * qj / (rij2 * rij);
                pot[7] += sQQ * R_FQQ * qi * qj / rij;
                qq.x = fij * rx; qq.y = fij * ry; qq.z = fij * rz;
            }
            double fx = lj.x + qq.x, fy = lj.y + qq.y, fz = lj.z + qq.z;
            vir[0] -= rx * fx; vir[1] -= rx * fy; vir[2] -= rx * fz;
            vir[3] -= ry * fx; vir[4] -= ry * fy; vir[5] -= ry * fz;
            vir[6] -= rz * fx; vir[7] -= rz * fy; vir[8] -= rz * fz;
            f[i] += -lj.x + -qq.x; f[N + i] += -lj.y + -qq.y; f[2 * N + i] += -lj.z + -qq.z;
            f[j] += lj.x + qq.x; f[N + j] += lj.y + qq.y; f[2 * N + j] += lj.z + qq.z;
        }
        /* angles: compute_angle_1 / compute_angle_5 (force.c:956-1052) */
        for (int t = rg[2 * M1 + m]; t < rg[2 * M1 + m + 1]; ++t) {
            int i = p->angle_atoms[3 * t], j = p->angle_atoms[3 * t + 1], k = p->angle_atoms[3 * t + 2];
            const double* ap = p->angle_par + 4 * t;
            double a0 = ap[0], ka = ap[1];
            ovec r21 = ov_pbc(o, r[i] - r[j], r[N + i] - r[N + j], r[2 * N + i] - r[2 * N + j]);
            double r21_d = sqrt(ov_dot(r21, r21));
            ovec u21 = ov_scale(1.0 / r21_d, r21);
            ovec r32 = ov_pbc(o, r[j] - r[k], r[N + j] - r[N + k], r[2 * N + j] - r[2 * N + k]);
            double r32_d = sqrt(ov_dot(r32, r32));
            ovec u32 = ov_scale(1.0 / r32_d, r32);
            double cos123 = -ov_dot(u21, u32);
            double theta123 = acos(cos123);
            double sin123 = sin(theta123);
            double delta_a = a0 * R_PI / 180.0 - theta123;
            double fij = ka * delta_a;
            ovec fi, fk, fj;
            fi.x = (cos123 * u21.x + u32.x) / (sin123 * r21_d) * fij;
            fi.y = (cos123 * u21.y + u32.y) / (sin123 * r21_d) * fij;
            fi.z = (cos123 * u21.z + u32.z) / (sin123 * r21_d) * fij;
            fk.x = -(cos123 * u32.x + u21.x) / (sin123 * r32_d) * fij;
            fk.y = -(cos123 * u32.y + u21.y) / (sin123 * r32_d) * fij;
            fk.z = -(cos123 * u32.z + u21.z) / (sin123 * r32_d) * fij;
            fj.x = -fi.x - fk.x; fj.y = -fi.y - fk.y; fj.z = -fi.z - fk.z;
            pot[3] += 0.5 * ka * delta_a * delta_a;
            ov_virial(r21, fi, vir);
            ovec m32 = {-r32.x, -r32.y, -r32.z};
            ov_virial(m32, fk, vir);
            if (p->angle_funct[t] == 5 && ap[3] > 0.0) {
                ovec bj = bond_term(o, r, i, k, 1, ap + 2, pot, vir);
                fi.x += -bj.x; fi.y += -bj.y; fi.z += -bj.z;
                fk.x += bj.x; fk.y += bj.y; fk.z += bj.z;
            }
            ov_add(f, N, i, fi);
            ov_add(f, N, j, fj);
            ov_add(f, N, k, fk);
        }
        /* dihedrals: compute_dihedral_3 (force.c:1100-1230) */
        for (int t = rg[3 * M1 + m]; t < rg[3 * M1 + m + 1]; ++t) {
            const int* q = p->dihedral_atoms + 4 * t;
            const double* c = p->dihedral_par + 6 * t;
            int i = q[0], j = q[1], k = q[2], l = q[3];
            ovec r21 = ov_pbc(o, r[i] - r[j], r[N + i] - r[N + j], r[2 * N + i] - r[2 * N + j]);
            double r21_d = sqrt(ov_dot(r21, r21));
            ovec u21 = ov_scale(1.0 / r21_d, r21);
            ovec r32 = ov_pbc(o, r[j] - r[k], r[N + j] - r[N + k], r[2 * N + j] - r[2 * N + k]);
            double r32_d = sqrt(ov_dot(r32, r32));
            ovec u32 = ov_scale(1.0 / r32_d, r32);
            ovec r43 = ov_pbc(o, r[k] - r[l], r[N + k] - r[N + l], r[2 * N + k] - r[2 * N + l]);
            double r43_d = sqrt(ov_dot(r43, r43));
            ovec u43 = ov_scale(1.0 / r43_d, r43);
            double cos123 = -ov_dot(u21, u32);
            double theta123 = acos(cos123);
            double sin123 = sin(theta123);
            double cos234 = -ov_dot(u32, u43);
            double theta234 = acos(cos234);
            double sin234 = sin(theta234);
            double cos_phi = (cos123 * cos234 - ov_dot(u21, u43)) / (sin123 * sin234);
            pot[4] += (c[0] - c[1] * cos_phi + c[2] * pow(cos_phi, 2.0) - c[3] * pow(cos_phi, 3.0) +
                       c[4] * pow(cos_phi, 4.0) - c[5] * pow(cos_phi, 5.0));
            double fij = -(-c[1] + c[2] * 2.0 * cos_phi - c[3] * 3.0 * pow(cos_phi, 2.0) +
                           c[4] * 4.0 * pow(cos_phi, 3.0) - c[5] * 5.0 * pow(cos_phi, 4.0));
            ovec a = ov_cross(u21, u32), b = ov_cross(u43, u32);
            double const_1 = ov_dot(u43, a) / (r21_d * pow(sin123, 3.0) * sin234);
            double const_4 = ov_dot(u43, a) / (r43_d * pow(sin234, 3.0) * sin123);
            double c123 = r21_d * cos123 / r32_d - 1.0;
            double b432 = r43_d * cos234 / r32_d;
            ovec fi = {-const_1 * a.x * fij, -const_1 * a.y * fij, -const_1 * a.z * fij};
            ovec fl = {-const_4 * b.x * fij, -const_4 * b.y * fij, -const_4 * b.z * fij};
            ovec fj = {c123 * fi.x - b432 * fl.x, c123 * fi.y - b432 * fl.y, c123 * fi.z - b432 * fl.z};
            ovec fk = {-fi.x - fj.x - fl.x, -fi.y - fj.y - fl.y, -fi.z - fj.z - fl.z};
            ovec rv = {r[i], r[N + i], r[2 * N + i]};
            ov_virial(rv, fi, vir);
            rv.x -= r21.x; rv.y -= r21.y; rv.z -= r21.z;
            ov_virial(rv, fj, vir);
            rv.x -= r32.x; rv.y -= r32.y; rv.z -= r32.z;
            ov_virial(rv, fk, vir);
            rv.x -= r43.x; rv.y -= r43.y; rv.z -= r43.z;
            ov_virial(rv, fl, vir);
            ov_add(f, N, i, fi);
            ov_add(f, N, j, fj);
            ov_add(f, N, k, fk);
            ov_add(f, N, l, fl);
        }
    }
}

/* build_vsites (force.c:1582-1629) */
void oref_vsite_build(oref* o, double* r) {
    const capac_params* p = &o->p;
    int N = o->N, M = p->n_mols;
    if (!p->mol_terms) return;
    const int* rg = p->mol_terms + 4 * (M + 1);
    for (int m = 0; m < M; ++m)
        for (int t = rg[m]; t < rg[m + 1]; ++t) {
            const int* q = p->vsite_atoms + 4 * t;
            int i = q[0], j = q[1], k = q[2], s = q[3];
            double a = p->vsite_par[3 * t], b = p->vsite_par[3 * t + 1], c = p->vsite_par[3 * t + 2];
            ovec rij = ov_pbc(o, r[j] - r[i], r[N + j] - r[N + i], r[2 * N + j] - r[2 * N + i]);
            ovec rik = ov_pbc(o, r[k] - r[i], r[N + k] - r[N + i], r[2 * N + k] - r[2 * N + i]);
            ovec x = ov_cross(rij, rik);
            r[s] = r[i] + a * rij.x + b * rik.x + c * x.x;
            r[N + s] = r[N + i] + a * rij.y + b * rik.y + c * x.y;
            r[2 * N + s] = r[2 * N + i] + a * rij.z + b * rik.z + c * x.z;
        }
}

/* mpi_vsites with compute_vsite_4 (force.c:1060-1090, 1636-1692) */
void oref_vsite_spread(oref* o, const double* r, double* f) {
    const capac_params* p = &o->p;
    int N = o->N, M = p->n_mols;
    if (!p->mol_terms) return;
    const int* rg = p->mol_terms + 4 * (M + 1);
    for (int m = 0; m < M; ++m)
        for (int t = rg[m]; t < rg[m + 1]; ++t) {
            const int* q = p->vsite_atoms + 4 * t;
            int i = q[0], j = q[1], k = q[2], s = q[3];
            double a = p->vsite_par[3 * t], b = p->vsite_par[3 * t + 1], c = p->vsite_par[3 * t + 2];
            double xij = pbc1(r[j] - r[i], o->box[0]), yij = pbc1(r[N + j] - r[N + i], o->box[1]);
            double zij = pbc1(r[2 * N + j] - r[2 * N + i], o->box[2]);
            double xik = pbc1(r[k] - r[i], o->box[0]), yik = pbc1(r[N + k] - r[N + i], o->box[1]);
            double zik = pbc1(r[2 * N + k] - r[2 * N + i], o->box[2]);
            double fxs = f[s], fys = f[N + s], fzs = f[2 * N + s];
            ovec fj, fk, fi;
            fj.x = a * fxs - c * zik * fys + c * yik * fzs;
            fj.y = c * zik * fxs + a * fys - c * xik * fzs;
            fj.z = -c * yik * fxs + c * xik * fys + a * fzs;
            fk.x = b * fxs + c * zij * fys - c * yij * fzs;
            fk.y = -c * zij * fxs + b * fys + c * xij * fzs;
            fk.z = c * yij * fxs - c * xij * fys + b * fzs;
            fi.x = fxs - fj.x - fk.x;
            fi.y = fys - fj.y - fk.y;
            fi.z = fzs - fj.z - fk.z;
            ov_add(f, N, i, fi);
            ov_add(f, N, j, fj);
            ov_add(f, N, k, fk);
            f[s] = 0.0; f[N + s] = 0.0; f[2 * N + s] = 0.0;
        }
}

/* ------------------------------------------------------------- Sutton-Chen */

/* mpi_qsc_dens + mpi_qsc_force (force.c:385-542) */
void oref_qsc(oref* o, const double* r, double* f, double* pot, double* vir, double* isd_out) {
    const capac_params* p = &o->p;
    int N = o->N, Nm = o->Nm;
    const double *rx = r, *ry = r + N, *rz = r + 2 * N;
    double t0 = now();
#ifdef ORACLE_OMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < Nm; ++i) {
        double dens = 0.0;
        for (int j = 0; j < Nm; ++j) {
            if (j == i) continue;
            double x = rx[j] - rx[i], y = ry[j] - ry[i], z = rz[j] - rz[i];
            pbc12(o->box, &x, &y, &z);
            double r2 = dist2(x, y, z);
            if (r2 < o->rc2) dens += pow((p->qsc_a / sqrt(r2)), p->qsc_m);
        }
        o->isd[i] = 1.0 / sqrt(dens);
    }
    double t1 = now();
    o->timers[1] += t1 - t0;
    double cm_2 = p->qsc_c * p->qsc_m / 2.0;
#ifdef ORACLE_OMP
    double p1 = 0.0, vl[9] = {0};
#pragma omp parallel for schedule(static) reduction(+ : p1, vl[:9])
    for (int i = 0; i < Nm; ++i) {
        double* potp = &p1;
        double* virp = vl;
#else
    for (int i = 0; i < Nm; ++i) {
        double* potp = &pot[1];
        double* virp = vir;
#endif
        for (int j = 0; j < Nm; ++j) {
            if (j == i) continue;
            double x = rx[j] - rx[i], y = ry[j] - ry[i], z = rz[j] - rz[i];
            pbc12(o->box, &x, &y, &z);
            double r2 = dist2(x, y, z);
            if (r2 < o->rc2) {
                double rij = sqrt(r2);
                double a_rij = p->qsc_a / rij;
                double a_n = pow(a_rij, p->qsc_n);
                double a_m = pow(a_rij, p->qsc_m);
                double sum = o->isd[i] + o->isd[j];
                double fij = p->qsc_eps * (p->qsc_n * a_n - cm_2 * sum * a_m) / r2;
                double fx = fij * x, fy = fij * y, fz = fij * z;
                f[i] -= fx;
                f[N + i] -= fy;
                f[2 * N + i] -= fz;
                /* virial_rvec_fvec_half (force.c:68-82) */
                virp[0] -= x * fx * 0.5; virp[1] -= x * fy * 0.5; virp[2] -= x * fz * 0.5;
                virp[3] -= y * fx * 0.5; virp[4] -= y * fy * 0.5; virp[5] -= y * fz * 0.5;
                virp[6] -= z * fx * 0.5; virp[7] -= z * fy * 0.5; virp[8] -= z * fz * 0.5;
                *potp += p->qsc_eps * a_n * 0.5;
            }
        }
        *potp -= p->qsc_eps * p->qsc_c / o->isd[i];
    }
#ifdef ORACLE_OMP
    pot[1] += p1;
    for (int k = 0; k < 9; ++k) vir[k] += vl[k];
#endif
    o->timers[2] += now() - t1;
    if (isd_out) memcpy(isd_out, o->isd, Nm * sizeof(double));
}

/* ---------------------------------------------------------------- CPIM RHS */

/* mpi_cpff_vec_ext (cpff.c:43-198) */
void oref_vec_ext(oref* o, const double* r, double* b) {
    const capac_params* p = &o->p;
    int N = o->N, Nm = o->Nm;
    const double *rx = r, *ry = r + N, *rz = r + 2 * N;
    double t0 = now();
#ifdef ORACLE_OMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < Nm; ++i) {
        double ex = 0.0, ey = 0.0, ez = 0.0, pm = 0.0;
        for (int a = 0; a < N; ++a) {
            if (a < Nm) continue;                       /* is_metal */
            double qi = p->charge[a];
            if (0.0 == qi) continue;
            double x = rx[i] - rx[a], y = ry[i] - ry[a], z = rz[i] - rz[a];
            pbc12(o->box, &x, &y, &z);
            double r2 = dist2(x, y, z);
            double e_field = 0.0, v_field = 0.0;
            if (r2 < o->rc2) {
                double rij = sqrt(r2);
                if (p->use_coulomb == 1) {
                    double erfc_arij = erfc(o->alpha * rij);
                    double ex2 = exp(-o->alpha * o->alpha * r2);
                    e_field += R_FQQ * qi * (erfc_arij / r2 + o->a2_sqrtPI * ex2 / rij - o->wolfConst) / rij;
                    v_field += R_FQQ * qi * (erfc_arij / rij - o->erfc_arCut + o->wolfConst * (rij - o->rc));
                } else {
                    e_field += R_FQQ * qi / (r2 * rij);
                    v_field += R_FQQ * qi / rij;
                }
            }
            double s = e_field * R_CONV_E;
            ex += s * x;
            ey += s * y;
            ez += s * z;
            pm += v_field * R_CONV_V;
        }
        b[i * 3 + 0] = ex;
        b[i * 3 + 1] = ey;
        b[i * 3 + 2] = ez;
        b[Nm * 3 + i] = pm;
        b[i * 3 + 0] += p->efield[0] * R_CONV_E;
        b[i * 3 + 1] += p->efield[1] * R_CONV_E;
        b[i * 3 + 2] += p->efield[2] * R_CONV_E;
        double pot_ext = p->efield[0] * rx[i] + p->efield[1] * ry[i] + p->efield[2] * rz[i];
        b[Nm * 3 + i] -= pot_ext * R_CONV_V;
    }
    for (int k = 0; k < o->n_np; ++k) b[Nm * 4 + k] = p->np_charge[k];
    o->timers[6] += now() - t0;
}

/* ------------------------------------------------------------ relay matrix */

typedef struct {
    long nnz, cap;
    long *row, *col;
    double* val;
} coo_buf;

static inline void coo_push(coo_buf* o, long row, long col, double v) {
    if (o->nnz >= o->cap) {
        long nc = o->cap ? 2 * o->cap : 1024;
        o->row = realloc(o->row, nc * sizeof(long));
        o->col = realloc(o->col, nc * sizeof(long));
        o->val = realloc(o->val, nc * sizeof(double));
        if (!o->row || !o->col || !o->val) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
        o->cap = nc;
    }
    o->row[o->nnz] = row;
    o->col[o->nnz] = col;
    o->val[o->nnz] = v;
    ++o->nnz;
}

/* the entries of metal row-group i of mpi_cpff_mat_relay_COO (cpff.c:245-478),
 * in the reference's push order */
static void relay_group(oref* o, const double* r, int i, coo_buf* out) {
    const capac_params* p = &o->p;
    int N = o->N, Nm = o->Nm;
    const double *rx = r, *ry = r + N, *rz = r + 2 * N;
    static const double delta[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    for (int j = 0; j < Nm; ++j) {
        double rv[3] = {0, 0, 0}, rij2 = 0, rij = 0;
        double c1 = 0, c2 = 0, c3 = 0;
        if (j != i) {
            double x = rx[j] - rx[i], y = ry[j] - ry[i], z = rz[j] - rz[i];
            pbc12(o->box, &x, &y, &z);
            x *= R_NM2BOHR; y *= R_NM2BOHR; z *= R_NM2BOHR;
            rij2 = dist2(x, y, z);
            rij = sqrt(rij2);
            rv[0] = x; rv[1] = y; rv[2] = z;
            double e_pp = exp(-(rij2 * o->inv_R_pp * o->inv_R_pp));
            double inv_Rpp3 = pow(o->inv_R_pp, 3.0);
            c1 = 1.0 / pow(rij, 5.0) *
                 (erf(rij * o->inv_R_pp) - 2.0 * R_INV_SQRT_PI * o->inv_R_pp * rij * e_pp);
            c2 = 4.0 * R_INV_SQRT_PI * inv_Rpp3 / rij2 * e_pp;
            double e_pq = exp(-(rij2 * o->inv_R_pq * o->inv_R_pq));
            c3 = 1.0 / pow(rij, 3.0) *
                 (erf(rij * o->inv_R_pq) - 2.0 * R_INV_SQRT_PI * o->inv_R_pq * rij * e_pq);
        }
        for (int a = 0; a < 3; ++a) {
            long irow = i * 3 + a;
            for (int b = 0; b < 3; ++b) {
                long icol = j * 3 + b;
                if (j == i) {
                    if (b == a) {
                        double el = o->inv_polar;
                        o->diag[irow] = el;
                        if (fabs(el) > R_THRSHD) coo_push(out, irow, icol, el);
                    }
                } else {
                    double el = -(c1 * (3.0 * rv[a] * rv[b] - delta[a][b] * rij2) - c2 * rv[a] * rv[b]);
                    if (fabs(el) > R_THRSHD) coo_push(out, irow, icol, el);
                }
            }
            if (j != i) {
                double el = c3 * rv[a];
                if (fabs(el) > R_THRSHD) coo_push(out, irow, (long)Nm * 3 + j, el);
            }
        }
        long irow = (long)Nm * 3 + i;
        for (int b = 0; b < 3; ++b) {
            if (j != i) {
                double el = -c3 * rv[b];
                if (fabs(el) > R_THRSHD) coo_push(out, irow, j * 3 + b, el);
            }
        }
        if (j == i) {
            double el = -o->inv_capac;
            o->diag[irow] = el;
            if (fabs(el) > R_THRSHD) coo_push(out, irow, (long)Nm * 3 + j, el);
        } else {
            double el = -erf(rij * o->inv_R_qq) / rij;
            if (fabs(el) > R_THRSHD) coo_push(out, irow, (long)Nm * 3 + j, el);
        }
        if (j == 0) {
            for (int k = 0; k < o->n_np; ++k)
                if (i >= p->np_first[k] && i <= p->np_last[k])
                    coo_push(out, irow, (long)Nm * 4 + k, 1.0);
        }
    }
}

/* mpi_cpff_mat_relay_COO (cpff.c:205-519) at one rank, same element order.
 * The OpenMP build gives each thread a contiguous range of row-groups
 * (static schedule) and concatenates the threads' entries in thread order,
 * which is the sequential order. */
long oref_relay(oref* o, const double* r) {
    const capac_params* p = &o->p;
    int Nm = o->Nm;
    double t0 = now();
    coo_buf main = {0, o->cap, o->row, o->col, o->val};
    if (!o->group_start) o->group_start = zalloc((size_t)Nm + 1, sizeof(long));
#ifdef ORACLE_OMP
    int nth = omp_get_max_threads();
    coo_buf* tb = zalloc(nth, sizeof(coo_buf));
    long* cnt = zalloc((size_t)Nm + 1, sizeof(long));
#pragma omp parallel
    {
        coo_buf* b = &tb[omp_get_thread_num()];
#pragma omp for schedule(static)
        for (int i = 0; i < Nm; ++i) {
            long before = b->nnz;
            relay_group(o, r, i, b);
            cnt[i] = b->nnz - before;
        }
    }
    long total = 0;
    for (int t = 0; t < nth; ++t) total += tb[t].nnz;
    if (total > main.cap) {
        main.row = realloc(main.row, total * sizeof(long));
        main.col = realloc(main.col, total * sizeof(long));
        main.val = realloc(main.val, total * sizeof(double));
        if (!main.row || !main.col || !main.val) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
        main.cap = total;
    }
    for (int t = 0; t < nth; ++t) {
        memcpy(main.row + main.nnz, tb[t].row, tb[t].nnz * sizeof(long));
        memcpy(main.col + main.nnz, tb[t].col, tb[t].nnz * sizeof(long));
        memcpy(main.val + main.nnz, tb[t].val, tb[t].nnz * sizeof(double));
        main.nnz += tb[t].nnz;
        free(tb[t].row); free(tb[t].col); free(tb[t].val);
    }
    free(tb);
    long acc = 0;
    for (int i = 0; i < Nm; ++i) { o->group_start[i] = acc; acc += cnt[i]; }
    o->group_start[Nm] = acc;
    free(cnt);
#else
    for (int i = 0; i < Nm; ++i) {
        o->group_start[i] = main.nnz;
        relay_group(o, r, i, &main);
    }
    o->group_start[Nm] = main.nnz;
#endif
    for (int k = 0; k < o->n_np; ++k)
        for (int i = 0; i < Nm; ++i)
            if (i >= p->np_first[k] && i <= p->np_last[k])
                coo_push(&main, (long)Nm * 4 + k, (long)Nm * 3 + i, 1.0);
    o->nnz = main.nnz; o->cap = main.cap;
    o->row = main.row; o->col = main.col; o->val = main.val;
    o->timers[5] += now() - t0;
    return o->nnz;
}

void oref_relay_coo(const oref* o, long* row, long* col, double* val, double* diag) {
    if (row) memcpy(row, o->row, o->nnz * sizeof(long));
    if (col) memcpy(col, o->col, o->nnz * sizeof(long));
    if (val) memcpy(val, o->val, o->nnz * sizeof(double));
    if (diag) memcpy(diag, o->diag, o->n_mat * sizeof(double));
}

void oref_relay_dense(const oref* o, double* dense) {
    long n = o->n_mat;
    memset(dense, 0, (size_t)n * n * sizeof(double));
    for (long k = 0; k < o->nnz; ++k) dense[o->row[k] * n + o->col[k]] += o->val[k];
}

/* ------------------------------------------------------------------ solver */

/* mpi_mat_vec_COO (bicg_stab.c:46-63).  A row-group's entries touch only its
 * own rows, so the OpenMP build runs row-groups in parallel, each over its
 * entries in COO order: every row sums in the sequential order. */
static void spmv_coo(const oref* o, const double* x, double* Ax) {
    for (long i = 0; i < o->n_mat; ++i) Ax[i] = 0.0;
#ifdef ORACLE_OMP
    const long* gs = o->group_start;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < o->Nm; ++g)
        for (long k = gs[g]; k < gs[g + 1]; ++k) Ax[o->row[k]] += o->val[k] * x[o->col[k]];
    for (long k = gs[o->Nm]; k < o->nnz; ++k) Ax[o->row[k]] += o->val[k] * x[o->col[k]];
#else
    for (long k = 0; k < o->nnz; ++k) Ax[o->row[k]] += o->val[k] * x[o->col[k]];
#endif
}

/* y = A x with the assembled COO relay (oref_relay first) */
void oref_relay_apply(const oref* o, const double* x, double* y) { spmv_coo(o, x, y); }

/* mpi_vec_vec (bicg_stab.c:121-178) at one rank */
static double vdot(const oref* o, const double* a, const double* b) {
    int Nm = o->Nm;
    double s = 0.0;
    for (int i = 0; i < Nm; ++i) {
        s += a[i * 3] * b[i * 3];
        s += a[i * 3 + 1] * b[i * 3 + 1];
        s += a[i * 3 + 2] * b[i * 3 + 2];
        s += a[Nm * 3 + i] * b[Nm * 3 + i];
    }
    for (int k = 0; k < o->n_np; ++k) s += a[Nm * 4 + k] * b[Nm * 4 + k];
    return s;
}

/* mpi_precon_bicg_stab_COO (bicg_stab.c:253-573); x in/out (warm start).
 * Returns CAPAC_SOLVE_*; the reference exits the program on FAILED. */
int oref_solve(oref* o, const double* b, double* x, int* iters, double* rTr_out, double* bTb_out) {
    int n = o->n_mat, Nm = o->Nm;
    double t0 = now();
    double rho_old = 1.0, omega = 1.0, alpha = 1.0, rho, beta;
    double bTb = vdot(o, b, b);
    int kind = CAPAC_SOLVE_CONVERGED, iter = 0;
    double rTr = 0.0;
    if (bTb < 1.0e-10) { kind = CAPAC_SOLVE_ZERO_RHS; goto out; }
    double criteria = bTb * o->solve_crit;
    spmv_coo(o, x, o->Ax);
    for (int i = 0; i < n; ++i) {
        o->r0[i] = b[i] - o->Ax[i];
        o->r[i] = o->r0[i];
        o->v[i] = 0.0;
        o->pp[i] = 0.0;
        o->K[i] = i < 4 * Nm ? 1.0 / o->diag[i] : 1.0;
    }
    for (;;) {
        rho = vdot(o, o->r0, o->r);
        if (rho < 1.0e-10 && iter == 0) { kind = CAPAC_SOLVE_ZERO_RHO; goto out; }
        beta = (rho / rho_old) * (alpha / omega);
        for (int i = 0; i < n; ++i) {
            o->pp[i] = o->r[i] + beta * (o->pp[i] - omega * o->v[i]);
            o->y[i] = i < 4 * Nm ? o->pp[i] * o->K[i] : o->pp[i];
        }
        spmv_coo(o, o->y, o->v);
        double r0_v = vdot(o, o->r0, o->v);
        alpha = rho / r0_v;
        for (int i = 0; i < n; ++i) {
            o->s[i] = o->r[i] - alpha * o->v[i];
            o->z[i] = i < 4 * Nm ? o->s[i] * o->K[i] : o->s[i];
        }
        spmv_coo(o, o->z, o->t);
        for (int i = 0; i < n; ++i) o->Kt[i] = i < 4 * Nm ? o->t[i] * o->K[i] : o->t[i];
        double Kt_Ks = vdot(o, o->Kt, o->z);
        double Kt_Kt = vdot(o, o->Kt, o->Kt);
        omega = Kt_Ks / Kt_Kt;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * o->y[i] + omega * o->z[i];
            o->r[i] = o->s[i] - omega * o->t[i];
        }
        rho_old = rho;
        iter++;
        rTr = vdot(o, o->r, o->r);
        if (rTr < criteria) break;
        if (iter > R_MAX_ITER) {
            kind = rTr < criteria * 100.0 ? CAPAC_SOLVE_LOOSENED : CAPAC_SOLVE_FAILED;
            break;
        }
    }
out:
    o->solve_iters = iter;
    o->solve_kind = kind;
    o->solve_rTr = rTr;
    o->solve_bTb = bTb;
    if (iters) *iters = iter;
    if (rTr_out) *rTr_out = rTr;
    if (bTb_out) *bTb_out = bTb;
    o->timers[7] += now() - t0;
    return kind;
}

/* -------------------------------------------------------------- CPIM force */

/* mpi_cpff_force (cpff.c:526-950) at one rank */
/* the terms of metal atom i in mpi_cpff_force (cpff.c:560-948): metal-water,
 * self, metal-metal and external field */
static void cpff_one(const oref* o, const double* r, const double* vec_pq, int i, double* f,
                     double* pot, double* vir) {
    const capac_params* p = &o->p;
    int N = o->N, Nm = o->Nm;
    const double *rx = r, *ry = r + N, *rz = r + 2 * N;
    static const double delta[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    double alpha = o->alpha;
    double alpha3 = pow(alpha, 3.0);
    double inv_Rpp3 = pow(o->inv_R_pp, 3.0);
    {
        for (int a_at = 0; a_at < N; ++a_at) {
            if (a_at < Nm) continue;
            if (0.0 == p->charge[a_at]) continue;
            double x = rx[i] - rx[a_at], y = ry[i] - ry[a_at], z = rz[i] - rz[a_at];
            pbc12(o->box, &x, &y, &z);
            double rij2 = dist2(x, y, z);
            if (rij2 < o->rc2) {
                double rij = sqrt(rij2);
                double qi = p->charge[a_at];
                int j = i;
                double qj = vec_pq[Nm * 3 + j];
                double rv[3] = {x, y, z};
                double fv[3] = {0., 0., 0.};
                double const_a0 = 0.0, erfc_arij = 0.0, exp2 = 0.0;
                if (p->use_coulomb == 1) {
                    erfc_arij = erfc(alpha * rij);
                    exp2 = exp(-alpha * alpha * rij2);
                    const_a0 = (erfc_arij / rij2 + o->a2_sqrtPI * exp2 / rij - o->wolfConst) / rij;
                    for (int a = 0; a < 3; ++a) fv[a] += const_a0 * rv[a] * R_FQQ * qi * qj;
                    pot[10] += R_FQQ * qi * qj * (erfc_arij / rij - o->erfc_arCut + o->wolfConst * (rij - o->rc));
                } else {
                    const_a0 = 1.0 / (rij2 * rij);
                    for (int a = 0; a < 3; ++a) fv[a] += const_a0 * rv[a] * R_FQQ * qi * qj;
                    pot[10] += R_FQQ * qi * qj / rij;
                }
                if (p->use_coulomb == 1) {
                    double const_a1 = 1.0 / pow(rij, 5.0) * (erfc_arij + o->a2_sqrtPI * rij * exp2);
                    double const_b1 = 4.0 * R_INV_SQRT_PI * alpha3 / rij2 * exp2;
                    double const_c1 = o->wolfConst / pow(rij, 3.0);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            fv[b] += -(const_a1 * (3.0 * rv[a] * rv[b] - delta[a][b] * rij2) +
                                       const_b1 * rv[a] * rv[b] -
                                       const_c1 * (rv[a] * rv[b] - delta[a][b] * rij2)) *
                                     R_FQQ * qi * vec_pq[j * 3 + a] / R_NM2BOHR;
                } else {
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            fv[b] += -((3.0 * rv[a] * rv[b] - delta[a][b] * rij2) / pow(rij, 5.0)) *
                                     R_FQQ * qi * vec_pq[j * 3 + a] / R_NM2BOHR;
                }
                for (int a = 0; a < 3; ++a)
                    pot[11] += -R_FQQ * qi * vec_pq[j * 3 + a] / R_NM2BOHR * const_a0 * rv[a];
                f[i] += fv[0];
                f[N + i] += fv[1];
                f[2 * N + i] += fv[2];
                f[a_at] -= fv[0];
                f[N + a_at] -= fv[1];
                f[2 * N + a_at] -= fv[2];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) vir[a * 3 + b] += -rv[a] * fv[b];
            }
        }
        /* self terms (cpff.c:727-742) */
        double qi = vec_pq[Nm * 3 + i];
        pot[12] += 0.5 * R_FQQ * qi * qi * (o->inv_capac * R_NM2BOHR);
        for (int a = 0; a < 3; ++a)
            pot[14] += 0.5 * R_FQQ * vec_pq[i * 3 + a] * vec_pq[i * 3 + a] * (o->inv_polar * R_NM2BOHR);

        /* metal-metal, all pairs, force on i only (cpff.c:762-927) */
        for (int j = 0; j < Nm; ++j) {
            if (j == i) continue;
            double x = rx[j] - rx[i], y = ry[j] - ry[i], z = rz[j] - rz[i];
            pbc12(o->box, &x, &y, &z);
            x *= R_NM2BOHR; y *= R_NM2BOHR; z *= R_NM2BOHR;
            double rij2 = dist2(x, y, z);
            double rij = sqrt(rij2);
            double qi2 = vec_pq[Nm * 3 + i];
            double qj = vec_pq[Nm * 3 + j];
            double rv[3] = {x, y, z};
            double fv[3] = {0., 0., 0.};
            double e_qq = exp(-(rij2 * o->inv_R_qq * o->inv_R_qq));
            double const_A0 = 1.0 / pow(rij, 3.0) *
                              (erf(rij * o->inv_R_qq) - 2.0 * R_INV_SQRT_PI * o->inv_R_qq * rij * e_qq);
            for (int a = 0; a < 3; ++a) fv[a] += const_A0 * rv[a] * R_FQQ_AU2 * qi2 * qj;
            pot[12] += R_FQQ_AU1 * qi2 * qj * erf(rij * o->inv_R_qq) / rij * 0.5;

            double e_pq = exp(-(rij2 * o->inv_R_pq * o->inv_R_pq));
            double inv_Rpq3 = pow(o->inv_R_pq, 3.0);
            double const_Apq0 = 1.0 / pow(rij, 3.0) *
                                (erf(rij * o->inv_R_pq) - 2.0 * R_INV_SQRT_PI * o->inv_R_pq * rij * e_pq);
            double const_A1 = 1.0 / pow(rij, 5.0) *
                              (erf(rij * o->inv_R_pq) - 2.0 * R_INV_SQRT_PI * o->inv_R_pq * rij * e_pq);
            double const_B1 = 4.0 * R_INV_SQRT_PI * inv_Rpq3 / rij2 * e_pq;
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) {
                    fv[b] += -(const_A1 * (3.0 * rv[a] * rv[b] - delta[a][b] * rij2) -
                               const_B1 * rv[a] * rv[b]) *
                             R_FQQ_AU2 * qi2 * vec_pq[j * 3 + a];
                    fv[b] += (const_A1 * (3.0 * rv[a] * rv[b] - delta[a][b] * rij2) -
                              const_B1 * rv[a] * rv[b]) *
                             R_FQQ_AU2 * qj * vec_pq[i * 3 + a];
                }
                pot[13] += -R_FQQ_AU1 * qi2 * vec_pq[j * 3 + a] * rv[a] * const_Apq0 * 0.5;
                pot[13] += R_FQQ_AU1 * qj * vec_pq[i * 3 + a] * rv[a] * const_Apq0 * 0.5;
            }
            double e_pp = exp(-(rij2 * o->inv_R_pp * o->inv_R_pp));
            double const_App1 = 1.0 / pow(rij, 5.0) *
                                (erf(rij * o->inv_R_pp) - 2.0 * R_INV_SQRT_PI * o->inv_R_pp * rij * e_pp);
            double const_Bpp1 = 4.0 * R_INV_SQRT_PI * inv_Rpp3 / rij2 * e_pp;
            double const_A2 = 1.0 / pow(rij, 7.0) *
                              (3.0 * erf(rij * o->inv_R_pp) -
                               6.0 * R_INV_SQRT_PI * o->inv_R_pp * rij * e_pp -
                               4.0 * pow(rij, 3.0) * R_INV_SQRT_PI * inv_Rpp3 * e_pp);
            double const_B2 = 8.0 * R_INV_SQRT_PI * pow(o->inv_R_pp, 5.0) / rij2 * e_pp;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    for (int c = 0; c < 3; ++c)
                        fv[c] += -(const_A2 * (5.0 * rv[a] * rv[b] * rv[c] -
                                               rij2 * (rv[a] * delta[b][c] + rv[b] * delta[c][a] +
                                                       rv[c] * delta[a][b])) -
                                   const_B2 * rv[a] * rv[b] * rv[c]) *
                                 R_FQQ_AU2 * vec_pq[i * 3 + a] * vec_pq[j * 3 + b];
                    pot[14] += -R_FQQ_AU1 * vec_pq[i * 3 + a] * vec_pq[j * 3 + b] *
                               ((3.0 * rv[a] * rv[b] - delta[a][b] * rij2) * const_App1 -
                                rv[a] * rv[b] * const_Bpp1) * 0.5;
                }
            f[i] -= fv[0];
            f[N + i] -= fv[1];
            f[2 * N + i] -= fv[2];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) vir[a * 3 + b] += -rv[a] * fv[b] * 0.5;
        }
        /* external field on the induced charge and dipole (cpff.c:930-948) */
        double q = vec_pq[Nm * 3 + i];
        f[i] += p->efield[0] * q;
        f[N + i] += p->efield[1] * q;
        f[2 * N + i] += p->efield[2] * q;
        pot[9] -= p->efield[0] * q * rx[i] + p->efield[1] * q * ry[i] + p->efield[2] * q * rz[i];
        double pv[3] = {vec_pq[i * 3], vec_pq[i * 3 + 1], vec_pq[i * 3 + 2]};
        pot[9] -= p->efield[0] * pv[0] + p->efield[1] * pv[1] + p->efield[2] * pv[2];
    }
}

/* mpi_cpff_force (cpff.c:526-950) at one rank.  The OpenMP build runs metal
 * atoms in parallel with per-thread force, energy and virial accumulators
 * (summation order changes: timing only). */
void oref_cpff_force(oref* o, const double* r, const double* vec_pq, double* f, double* pot,
                     double* vir) {
    int Nm = o->Nm;
    double t0 = now();
#ifdef ORACLE_OMP
    int N = o->N, nth = omp_get_max_threads();
    double* fb = zalloc((size_t)nth * 3 * N, sizeof(double));
    double* pb = zalloc((size_t)nth * 24, sizeof(double));
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
#pragma omp for schedule(dynamic, 8)
        for (int i = 0; i < Nm; ++i)
            cpff_one(o, r, vec_pq, i, fb + (size_t)tid * 3 * N, pb + (size_t)tid * 24, pb + (size_t)tid * 24 + 15);
    }
    for (int t = 0; t < nth; ++t) {
        for (size_t k = 0; k < 3 * (size_t)N; ++k) f[k] += fb[(size_t)t * 3 * N + k];
        for (int k = 9; k < 15; ++k) pot[k] += pb[t * 24 + k];
        for (int k = 0; k < 9; ++k) vir[k] += pb[t * 24 + 15 + k];
    }
    free(fb); free(pb);
#else
    for (int i = 0; i < Nm; ++i) cpff_one(o, r, vec_pq, i, f, pot, vir);
#endif
    o->timers[8] += now() - t0;
}

/* ------------------------------------------------------------- force call */

/* mpi_force (force.c:2097-2320) at one rank */
int oref_force(oref* o) {
    const capac_params* p = &o->p;
    int N = o->N;
    double t0 = now();
    oref_vsite_build(o, o->rr);
    memset(o->ff, 0, 3 * (size_t)N * sizeof(double));
    memset(o->pot, 0, sizeof o->pot);
    memset(o->vir, 0, sizeof o->vir);
    oref_bonded(o, o->rr, o->ff, o->pot, o->vir);
    oref_nonb(o, o->rr, o->ff, o->pot, o->vir);
    int rc = 0;
    if (o->Nm > 0) {
        oref_qsc(o, o->rr, o->ff, o->pot, o->vir, NULL);
        if (p->use_cpff) {
            oref_vec_ext(o, o->rr, o->vec_ext);
            oref_relay(o, o->rr);
            int kind = oref_solve(o, o->vec_ext, o->vec_pq, NULL, NULL, NULL);
            if (kind == CAPAC_SOLVE_FAILED) rc = CAPAC_ERR_SOLVER;
            oref_cpff_force(o, o->rr, o->vec_pq, o->ff, o->pot, o->vir);
        }
    }
    for (int i = 0; i < N; ++i) {
        double qi = p->charge[i];
        o->ff[i] += p->efield[0] * qi;
        o->ff[N + i] += p->efield[1] * qi;
        o->ff[2 * N + i] += p->efield[2] * qi;
        o->pot[9] -= p->efield[0] * qi * o->rr[i] + p->efield[1] * qi * o->rr[N + i] +
                     p->efield[2] * qi * o->rr[2 * N + i];
    }
    oref_vsite_spread(o, o->rr, o->ff);
    o->timers[0] += now() - t0;
    return rc;
}

void oref_get_force_out(const oref* o, double* f, double* pot, double* vir, double* vec_ext) {
    if (f) memcpy(f, o->ff, 3 * (size_t)o->N * sizeof(double));
    if (pot) memcpy(pot, o->pot, sizeof o->pot);
    if (vir) memcpy(vir, o->vir, sizeof o->vir);
    if (vec_ext && o->n_mat) memcpy(vec_ext, o->vec_ext, o->n_mat * sizeof(double));
}

void oref_solve_stats(const oref* o, int* iters, double* rTr, double* bTb, int* kind) {
    if (iters) *iters = o->solve_iters;
    if (rTr) *rTr = o->solve_rTr;
    if (bTb) *bTb = o->solve_bTb;
    if (kind) *kind = o->solve_kind;
}

/* -------------------------------------------------------------------- MD */

/* kinetic_energy (thermostat.c:34-63) */
static void kinetic(oref* o) {
    int N = o->N;
    double xx = 0.0, yy = 0.0, zz = 0.0;
    for (int i = 0; i < N; ++i) {
        double m = o->p.mass[i];
        if (m > 0.0) {
            xx += m * o->vv[i] * o->vv[i];
            yy += m * o->vv[N + i] * o->vv[N + i];
            zz += m * o->vv[2 * N + i] * o->vv[2 * N + i];
        }
    }
    o->eKsum = xx + yy + zz;
    o->inst_temp = o->eKsum / (R_K_BOLTZ * o->p.ndf);
}

/* remove_comm (thermostat.c:69-108) */
static void remove_com(oref* o) {
    int N = o->N;
    double vx = 0.0, vy = 0.0, vz = 0.0, ms = 0.0;
    for (int i = 0; i < N; ++i) {
        double m = o->p.mass[i];
        if (m > 0.0) {
            vx += o->vv[i] * m;
            vy += o->vv[N + i] * m;
            vz += o->vv[2 * N + i] * m;
            ms += m;
        }
    }
    vx /= ms; vy /= ms; vz /= ms;
    for (int i = 0; i < N; ++i)
        if (o->p.mass[i] > 0.0) {
            o->vv[i] -= vx;
            o->vv[N + i] -= vy;
            o->vv[2 * N + i] -= vz;
        }
}

/* nose_hoover_chain (thermostat.c:117-191) */
static void nhc(oref* o) {
    const int Nc = 5, Nys = 5, M = CAPAC_NHC;
    double w1 = 0.41449077179437571194;
    double w3 = -0.65796308717750284778;
    double w[5] = {w1, w1, w3, w1, w1};
    double qMass = o->qMass, eKsum = o->eKsum;
    double* aQ = o->aQ;
    double* vQ = o->vQ;
    double dt = o->p.dt, kT = o->kT, NkT = kT * o->p.ndf;
    double aa, scale = 1.0;
    aQ[0] = (eKsum - NkT) / qMass;
    for (int im = 0; im < M - 1; ++im) aQ[im + 1] = (qMass * vQ[im] * vQ[im] - kT) / qMass;
    for (int ic = 0; ic < Nc; ++ic)
        for (int iys = 0; iys < Nys; ++iys) {
            vQ[M - 1] += aQ[M - 1] * w[iys] * dt / (4 * Nc);
            for (int im = 0; im < M - 1; ++im) {
                aa = exp(-w[iys] * dt / (8 * Nc) * vQ[M - (im + 1)]);
                vQ[M - 1 - (im + 1)] = vQ[M - 1 - (im + 1)] * aa * aa +
                                       w[iys] * dt / (4 * Nc) * aQ[M - 1 - (im + 1)] * aa;
            }
            aa = exp(-w[iys] * dt / (2 * Nc) * vQ[0]);
            scale *= aa;
            aQ[0] = (eKsum * scale * scale - NkT) / qMass;
            for (int im = 0; im < M - 1; ++im) {
                aa = exp(-w[iys] * dt / (8 * Nc) * vQ[im + 1]);
                vQ[im] = vQ[im] * aa * aa + w[iys] * dt / (4 * Nc) * aQ[im] * aa;
                aQ[im + 1] = (qMass * vQ[im] * vQ[im] - kT) / qMass;
            }
            vQ[M - 1] += aQ[M - 1] * w[iys] * dt / (4 * Nc);
        }
    for (size_t i = 0; i < 3 * (size_t)o->N; ++i) o->vv[i] *= scale;
    o->eKsum *= scale * scale;
    o->inst_temp *= scale * scale;
}

/* rattle_1st (force.c:1907-1997); returns 0 or CAPAC_ERR_RATTLE1 */
static int rattle1(oref* o) {
    const capac_params* p = &o->p;
    int N = o->N;
    double tol2 = 2.0e-05, dt = p->dt;
    double *rx = o->rr, *ry = o->rr + N, *rz = o->rr + 2 * N;
    double *vx = o->vv, *vy = o->vv + N, *vz = o->vv + 2 * N;
    double *ox = o->old_r, *oy = o->old_r + N, *oz = o->old_r + 2 * N;
    for (int im = 0; im < p->n_mols; ++im) {
        int iter = 0, done = 0;
        while (0 == done) {
            done = 1;
            for (int c = p->mol_cstr[im]; c < p->mol_cstr[im + 1]; ++c) {
                int a = p->cstr_i[c], b = p->cstr_j[c];
                double rab = p->cstr_r[c];
                double pxab = pbc1(rx[a] - rx[b], o->box[0]);
                double pyab = pbc1(ry[a] - ry[b], o->box[1]);
                double pzab = pbc1(rz[a] - rz[b], o->box[2]);
                double pabSQ = pxab * pxab + pyab * pyab + pzab * pzab;
                double rabSQ = rab * rab;
                double diffSQ = rabSQ - pabSQ;
                if (fabs(diffSQ) > rabSQ * tol2) {
                    double rxab = pbc1(ox[a] - ox[b], o->box[0]);
                    double ryab = pbc1(oy[a] - oy[b], o->box[1]);
                    double rzab = pbc1(oz[a] - oz[b], o->box[2]);
                    double rpab = rxab * pxab + ryab * pyab + rzab * pzab;
                    double rma = 1.0 / p->mass[a], rmb = 1.0 / p->mass[b];
                    double gab = diffSQ / (2.0 * (rma + rmb) * rpab);
                    double dx = rxab * gab, dy = ryab * gab, dz = rzab * gab;
                    rx[a] += rma * dx; ry[a] += rma * dy; rz[a] += rma * dz;
                    rx[b] -= rmb * dx; ry[b] -= rmb * dy; rz[b] -= rmb * dz;
                    dx /= dt; dy /= dt; dz /= dt;
                    vx[a] += rma * dx; vy[a] += rma * dy; vz[a] += rma * dz;
                    vx[b] -= rmb * dx; vy[b] -= rmb * dy; vz[b] -= rmb * dz;
                    done = 0;
                }
            }
            iter++;
            if (iter > R_MAX_ITER) break;
        }
        if (0 == done) return CAPAC_ERR_RATTLE1;
    }
    return 0;
}

/* rattle_2nd (force.c:2004-2090), including its virial update */
static int rattle2(oref* o) {
    const capac_params* p = &o->p;
    int N = o->N;
    double tol = 1.0e-05, half_dt = 0.5 * p->dt;
    double *rx = o->rr, *ry = o->rr + N, *rz = o->rr + 2 * N;
    double *vx = o->vv, *vy = o->vv + N, *vz = o->vv + 2 * N;
    for (int im = 0; im < p->n_mols; ++im) {
        int iter = 0, done = 0;
        while (done == 0) {
            done = 1;
            for (int c = p->mol_cstr[im]; c < p->mol_cstr[im + 1]; ++c) {
                int a = p->cstr_i[c], b = p->cstr_j[c];
                double rab = p->cstr_r[c];
                double vxab = vx[a] - vx[b], vyab = vy[a] - vy[b], vzab = vz[a] - vz[b];
                double rxab = pbc1(rx[a] - rx[b], o->box[0]);
                double ryab = pbc1(ry[a] - ry[b], o->box[1]);
                double rzab = pbc1(rz[a] - rz[b], o->box[2]);
                double rvab = rxab * vxab + ryab * vyab + rzab * vzab;
                double rma = 1.0 / p->mass[a], rmb = 1.0 / p->mass[b];
                double rabSQ = rab * rab;
                double gab = -rvab / ((rma + rmb) * rabSQ);
                if (fabs(gab) > tol) {
                    double dx = rxab * gab, dy = ryab * gab, dz = rzab * gab;
                    vx[a] += rma * dx; vy[a] += rma * dy; vz[a] += rma * dz;
                    vx[b] -= rmb * dx; vy[b] -= rmb * dy; vz[b] -= rmb * dz;
                    dx /= half_dt; dy /= half_dt; dz /= half_dt;
                    o->vir[0] -= rxab * dx; o->vir[1] -= rxab * dy; o->vir[2] -= rxab * dz;
                    o->vir[3] -= ryab * dx; o->vir[4] -= ryab * dy; o->vir[5] -= ryab * dz;
                    o->vir[6] -= rzab * dx; o->vir[7] -= rzab * dy; o->vir[8] -= rzab * dz;
                    done = 0;
                }
            }
            iter++;
            if (iter > R_MAX_ITER) break;
        }
        if (done == 0) return CAPAC_ERR_RATTLE2;
    }
    return 0;
}

/* apply_pbc (force.c:127-165) */
static void wrap_molecules(oref* o) {
    const capac_params* p = &o->p;
    int N = o->N;
    double* rc[3] = {o->rr, o->rr + N, o->rr + 2 * N};
    double half[3] = {0.5 * o->box[0], 0.5 * o->box[1], 0.5 * o->box[2]};
    for (int im = 0; im < p->n_mols; ++im) {
        int first = p->mol_first[im], last = p->mol_last[im];
        int mid = (first + last) / 2;
        for (int c = 0; c < 3; ++c) {
            if (rc[c][mid] > o->box[c]) rc[c][mid] -= o->box[c];
            else if (rc[c][mid] < 0.0) rc[c][mid] += o->box[c];
        }
        for (int i = first; i <= last; ++i) {
            if (i == mid) continue;
            for (int c = 0; c < 3; ++c) {
                if (rc[c][i] - rc[c][mid] > half[c]) rc[c][i] -= o->box[c];
                else if (rc[c][i] - rc[c][mid] < -half[c]) rc[c][i] += o->box[c];
            }
        }
    }
}

/* get_fmax_rms (force.c:361-378) */
static void fmax_rms(oref* o) {
    int N = o->N;
    double fmax2 = 0.0, frms = 0.0;
    for (int i = 0; i < N; ++i) {
        double f2 = o->ff[i] * o->ff[i] + o->ff[N + i] * o->ff[N + i] + o->ff[2 * N + i] * o->ff[2 * N + i];
        frms += f2;
        if (f2 > fmax2) fmax2 = f2;
    }
    o->fmax = sqrt(fmax2);
    o->frms = sqrt(frms / N);
}

static void sum_pot(double* pot) {
    pot[0] = 0.0;
    for (int i = 1; i < 15; ++i) pot[0] += pot[i];
}

/* after the first force call (mpi-main.c:629-640) */
void oref_md_begin(oref* o) {
    sum_pot(o->pot);
    remove_com(o);
    kinetic(o);
    o->first_temp = o->inst_temp;
    o->ext_temp = o->p.ref_temp;
    fmax_rms(o);
}

/* one MD step (mpi-main.c:874-1013) */
int oref_md_step(oref* o, int step, int n_save) {
    const capac_params* p = &o->p;
    int N = o->N, Nm = o->Nm;
    if (step < p->n_heating)
        o->ref_temp = o->first_temp + (o->ext_temp - o->first_temp) * step / p->n_heating;
    else
        o->ref_temp = o->ext_temp;
    o->kT = R_K_BOLTZ * o->ref_temp;
    if (p->nvt) nhc(o);
    double dt = p->dt, dt_2 = 0.5 * p->dt;
    for (int i = 0; i < N; ++i) {
        if (p->fix_metal == 1 && i < Nm) continue;
        double im = p->mass[i] > 0.0 ? 1.0 / p->mass[i] : 0.0;
        o->vv[i] += dt_2 * o->ff[i] * im;
        o->vv[N + i] += dt_2 * o->ff[N + i] * im;
        o->vv[2 * N + i] += dt_2 * o->ff[2 * N + i] * im;
    }
    for (int i = 0; i < N; ++i) {
        if (p->fix_metal == 1 && i < Nm) continue;
        for (int c = 0; c < 3; ++c) {
            o->old_r[c * N + i] = o->rr[c * N + i];
            o->rr[c * N + i] += dt * o->vv[c * N + i];
        }
    }
    int rc = rattle1(o);
    if (rc) return rc;
    rc = oref_force(o);
    if (rc) return rc;
    sum_pot(o->pot);
    for (int i = 0; i < N; ++i) {
        if (p->fix_metal == 1 && i < Nm) continue;
        double im = p->mass[i] > 0.0 ? 1.0 / p->mass[i] : 0.0;
        o->vv[i] += dt_2 * o->ff[i] * im;
        o->vv[N + i] += dt_2 * o->ff[N + i] * im;
        o->vv[2 * N + i] += dt_2 * o->ff[2 * N + i] * im;
    }
    rc = rattle2(o);
    if (rc) return rc;
    kinetic(o);
    if (p->nvt) nhc(o);
    remove_com(o);
    kinetic(o);
    if (n_save > 0 && step % n_save == 0) {
        wrap_molecules(o);
        fmax_rms(o);
    }
    return 0;
}

void oref_md_observables(const oref* o, double* eksum, double* temp, double* ref_temp, double* vq,
                         double* fmax, double* frms, double* pot, double* vir) {
    if (eksum) *eksum = o->eKsum;
    if (temp) *temp = o->inst_temp;
    if (ref_temp) *ref_temp = o->ref_temp;
    if (vq) memcpy(vq, o->vQ, sizeof o->vQ);
    if (fmax) *fmax = o->fmax;
    if (frms) *frms = o->frms;
    if (pot) memcpy(pot, o->pot, sizeof o->pot);
    if (vir) memcpy(vir, o->vir, sizeof o->vir);
}

void oref_set_thermostat(oref* o, const double* vq) {
    if (vq) memcpy(o->vQ, vq, sizeof o->vQ);
}

void oref_timers(const oref* o, double* t15) { memcpy(t15, o->timers, sizeof o->timers); }

/* Stop test of the solve: rTr < crit * bTb; the reference's is 1e-8
 * (bicg_stab.c:300).  Tighter values serve comparisons that should not see
 * the solver tolerance. */
void oref_set_solve_crit(oref* o, double crit) { o->solve_crit = crit; }

/* ---------------------------------------------------- energy minimisation */

/* run_type em / cg (mpi-main.c:706-865), after oref_force and oref_md_begin:
 * s = F, F_old = F (mpi-main.c:716-726) */
void oref_em_begin(oref* o) {
    size_t n3 = 3 * (size_t)o->N;
    if (!o->em_s) {
        o->em_s = zalloc(n3, sizeof(double));
        o->em_old_f = zalloc(n3, sizeof(double));
    }
    memcpy(o->em_s, o->ff, n3 * sizeof(double));
    memcpy(o->em_old_f, o->ff, n3 * sizeof(double));
}

/* One minimiser step (mpi-main.c:730-776): move along s by length / F_max,
 * RATTLE_1, zero velocities, force call, energies summed, F_max and F_rms.
 * Returns 0, CAPAC_ERR_RATTLE1 or CAPAC_ERR_SOLVER. */
int oref_em_step(oref* o, double length, double* pot, double* fmax, double* frms) {
    const capac_params* p = &o->p;
    int N = o->N, Nm = o->Nm;
    for (int i = 0; i < N; ++i) {
        for (int c = 0; c < 3; ++c) o->old_r[c * N + i] = o->rr[c * N + i];
        if (p->fix_metal == 1 && i < Nm) continue;
        for (int c = 0; c < 3; ++c) o->rr[c * N + i] += o->em_s[c * N + i] / o->fmax * length;
    }
    int rc = rattle1(o);
    if (rc) return rc;
    memset(o->vv, 0, 3 * (size_t)N * sizeof(double));
    rc = oref_force(o);
    sum_pot(o->pot);
    fmax_rms(o);
    if (pot) memcpy(pot, o->pot, sizeof o->pot);
    if (fmax) *fmax = o->fmax;
    if (frms) *frms = o->frms;
    return rc;
}

/* New direction when not converged (mpi-main.c:821-849): s = F + gamma s,
 * gamma = 0 (em) or (F.F - F_old.F) / (F_old.F_old) (cg), F_old = F. */
void oref_em_direction(oref* o, int cg) {
    int N = o->N;
    const double *fx = o->ff, *fy = o->ff + N, *fz = o->ff + 2 * N;
    double *ox = o->em_old_f, *oy = o->em_old_f + N, *oz = o->em_old_f + 2 * N;
    double gamma = 0.0;
    if (cg) {
        double g22 = 0.0, g12 = 0.0, g11 = 0.0;
        for (int i = 0; i < N; ++i) {
            g22 += fx[i] * fx[i] + fy[i] * fy[i] + fz[i] * fz[i];
            g12 += ox[i] * fx[i] + oy[i] * fy[i] + oz[i] * fz[i];
            g11 += ox[i] * ox[i] + oy[i] * oy[i] + oz[i] * oz[i];
        }
        gamma = (g22 - g12) / g11;
    }
    for (size_t k = 0; k < 3 * (size_t)N; ++k) {
        o->em_s[k] = o->ff[k] + gamma * o->em_s[k];
        o->em_old_f[k] = o->ff[k];
    }
}
