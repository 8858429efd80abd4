// C ABI of libcapac_gpu.so (include/capac_gpu.h): context set-up, the force
// call that replaces the reference's mpi_force (force.c:2097-2320), the
// sub-operator entry points and the device MD loop (mpi-main.c:874-1013).
#include "capac_internal.h"
#include "relay_tile.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static thread_local char g_err[1024] = "";

void capac_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char* capac_last_error(void) { return g_err; }
extern "C" const char* capac_version(void) { return "capac_gpu 0.1 (gfx950, FP64 + FP32-refined CPIM)"; }

int pair_blocks(int N);
int pair_partial_blocks(int N);

namespace {

struct ReduceJobs {
    int part_off[SEG_COUNT], nblk[SEG_COUNT], nval[SEG_COUNT], red_off[SEG_COUNT];
};

// one workgroup per queued reduction: fixed-order sums of a segment's
// per-block partials [nblk x 16] into red
__global__ void k_reduce(const double* __restrict__ partial_all, ReduceJobs jobs, double* __restrict__ red) {
    __shared__ double s_w[16][16];
    const int job = blockIdx.x;
    reduce_rows_block(partial_all + (size_t)jobs.part_off[job] * 16, jobs.nblk[job], jobs.nval[job],
                      red + jobs.red_off[job], s_w, NULL);
}

__global__ void k_fill_d(double* __restrict__ p, size_t n, double v) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ void k_fill_i(int* __restrict__ p, size_t n, int v) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ void k_copy_d(const double* __restrict__ a, double* __restrict__ b, size_t n) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = a[k];
}

}  // namespace

double* partial_seg(capac_ctx* c, int seg) { return c->partial + (size_t)c->seg_off[seg] * 16; }

int flush_reductions(capac_ctx* c) {
    if (c->n_pending == 0) return 0;
    ReduceJobs jobs;
    for (int k = 0; k < c->n_pending; ++k) {
        jobs.part_off[k] = c->seg_off[c->pend_seg[k]];
        jobs.nblk[k] = c->pend_nblk[k];
        jobs.nval[k] = c->pend_nval[k];
        jobs.red_off[k] = c->pend_red[k];
    }
    k_reduce<<<c->n_pending, 1024, 0, c->stream>>>(c->partial, jobs, c->red);
    c->n_pending = 0;
    CHECK_HIP(hipGetLastError());
    return 0;
}

int reduce_energies(capac_ctx* c, int seg, int nblk, int nval, int off) {
    return c->need_energy ? reduce_partials(c, seg, nblk, nval, off) : 0;
}

int reduce_partials(capac_ctx* c, int seg, int nblk, int nval, int off) {
    if (seg < 0 || seg >= SEG_COUNT || nblk > c->seg_cap[seg] || nval > 16) {
        capac_set_error("internal: reduction of %d blocks exceeds segment %d", nblk, seg);
        return CAPAC_ERR_HIP;
    }
    // a segment still queued would be overwritten by its next producer
    for (int k = 0; k < c->n_pending; ++k)
        if (c->pend_seg[k] == seg) {
            int rc = flush_reductions(c);
            if (rc) return rc;
            break;
        }
    const int k = c->n_pending++;
    c->pend_seg[k] = seg;
    c->pend_nblk[k] = nblk;
    c->pend_nval[k] = nval;
    c->pend_red[k] = off;
    return 0;
}

static int fill_d(capac_ctx* c, double* p, size_t n, double v) {
    if (n == 0) return 0;
    k_fill_d<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(p, n, v);
    CHECK_HIP(hipGetLastError());
    return 0;
}

// zeroes red[off..off+n) after launching any queued reduction, so a queued
// result cannot land after the clear
static int clear_red(capac_ctx* c, int off, int n) {
    int rc = flush_reductions(c);
    return rc ? rc : fill_d(c, c->red + off, (size_t)n, 0.0);
}

static int fill_i(capac_ctx* c, int* p, size_t n, int v) {
    if (n == 0) return 0;
    k_fill_i<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(p, n, v);
    CHECK_HIP(hipGetLastError());
    return 0;
}

int fill_zero_i(capac_ctx* c, int* p, size_t n) { return fill_i(c, p, n, 0); }
int fill_zero_d(capac_ctx* c, double* p, size_t n) { return fill_d(c, p, n, 0.0); }

static int copy_d(capac_ctx* c, const double* a, double* b, size_t n) {
    if (n == 0) return 0;
    k_copy_d<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(a, b, n);
    CHECK_HIP(hipGetLastError());
    return 0;
}

int zero_outputs(capac_ctx* c) {
    int rc = clear_red(c, 0, 256);
    if (!rc) rc = fill_d(c, c->f, 3 * (size_t)c->bc.N, 0.0);
    return rc;
}

// ------------------------------------------------------------------ set-up

template <typename T>
static int upload(T** dst, const T* src, size_t n) {
    size_t bytes = (n ? n : 1) * sizeof(T);
    CHECK_HIP(hipMalloc(dst, bytes));
    if (src && n) CHECK_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    else CHECK_HIP(hipMemset(*dst, 0, bytes));
    return 0;
}

// (re)allocates the block-Jacobi inverses (none when bj_every is 0); the
// next assembly forms them
static int bj_alloc(capac_ctx* c) { return bj_alloc_dev(c); }

// (re)allocates the FP32 relay in `layout` (AUTO: tiles from
// CAPAC_TILES_MIN_METAL metal atoms on), zeroed: ROWS padded with zero
// coefficient rows up to a multiple of kRelay32RowPad (the SpMV reads whole
// row blocks), TILES with its partial-sum buffer.  The contents are formed by
// the next assembly.
static int relay32_alloc(capac_ctx* c, int layout) {
    const int Nm = c->bc.Nm;
    if (layout == CAPAC_LAYOUT_AUTO) layout = Nm >= CAPAC_TILES_MIN_METAL ? CAPAC_LAYOUT_TILES : CAPAC_LAYOUT_ROWS;
    if (c->relay32) { (void)hipFree(c->relay32); c->relay32 = NULL; }
    if (c->relay_part) { (void)hipFree(c->relay_part); c->relay_part = NULL; }
    if (c->relay_cnt) { (void)hipFree(c->relay_cnt); c->relay_cnt = NULL; }
    if (int rc = bj_alloc(c)) return rc;
    c->relay_layout = layout;
    c->relay_nb = tile_blocks(Nm);
    if (layout == CAPAC_LAYOUT_TILES) {
        c->relay_bytes = (size_t)tile_count(c->relay_nb) * kTileSlots * 4 * sizeof(float);
        const size_t part = (size_t)c->relay_nb * c->relay_nb * 4 * kTileEdge * sizeof(double);
        CHECK_HIP(hipMalloc(&c->relay_part, part));
        CHECK_HIP(hipMemset(c->relay_part, 0, part));
        CHECK_HIP(hipMalloc(&c->relay_cnt, sizeof(int) * c->relay_nb));
        CHECK_HIP(hipMemset(c->relay_cnt, 0, sizeof(int) * c->relay_nb));
    } else {
        const size_t rows = ((size_t)Nm + kRelay32RowPad - 1) / kRelay32RowPad * kRelay32RowPad;
        c->relay_bytes = rows * kRelay32Planes * c->relay_ld * sizeof(float);
    }
    CHECK_HIP(hipMalloc(&c->relay32, c->relay_bytes));
    CHECK_HIP(hipMemset(c->relay32, 0, c->relay_bytes));
    return 0;
}

static int alloc_nbr(capac_ctx* c, int stride) {
    if (c->nbr) (void)hipFree(c->nbr);
    c->nbr = NULL;
    c->nb_stride = stride;
    CHECK_HIP(hipMalloc(&c->nbr, (size_t)c->bc.N * stride * sizeof(int)));
    return 0;
}

static int alloc_cells(capac_ctx* c) {
    int* arrs[4] = {c->cell_count, c->cell_start, c->cell_fill, NULL};
    for (int k = 0; k < 3; ++k)
        if (arrs[k]) (void)hipFree(arrs[k]);
    CHECK_HIP(hipMalloc(&c->cell_count, sizeof(int) * (c->n_cell + 1)));
    CHECK_HIP(hipMalloc(&c->cell_start, sizeof(int) * (c->n_cell + 1)));
    CHECK_HIP(hipMalloc(&c->cell_fill, sizeof(int) * (c->n_cell + 1)));
    return 0;
}

static void derive_constants(capac_ctx* c, const capac_params* p) {
    BoxConst& b = c->bc;
    for (int d = 0; d < 3; ++d) { b.box[d] = p->box[d]; b.half[d] = p->box[d] * 0.5; }
    // the reference's derived constants, same expressions (mpi-main.c:170-178,
    // 202-203, 315-323)
    b.rc = p->rcut;
    b.rc2 = p->rcut * p->rcut;
    b.alpha = p->dsf_alpha;
    b.a2_sqrtPI = b.alpha * 2.0 * K_INV_SQRT_PI;
    b.wolfConst = erfc(b.alpha * b.rc) / b.rc2 + b.a2_sqrtPI * exp(-b.alpha * b.alpha * b.rc2) / b.rc;
    b.erfc_arCut = erfc(b.alpha * b.rc) / b.rc;
    b.inv_rc12 = 1.0 / pow(b.rc, 12.0);
    b.inv_rc6 = 1.0 / pow(b.rc, 6.0);
    b.use_coulomb = p->use_coulomb;
    b.use_vdw = p->use_vdw;
    b.qsc_n = p->qsc_n; b.qsc_m = p->qsc_m;
    b.qsc_eps = p->qsc_eps; b.qsc_c = p->qsc_c; b.qsc_a = p->qsc_a;
    b.inv_polar = 1.0 / p->cpff_polar;
    b.inv_capac = 1.0 / p->cpff_capac;
    const double pi = 3.14159265358979323846;
    double R_q = sqrt(2.0 / pi) * p->cpff_capac;
    double R_p = pow(sqrt(2.0 / pi) * p->cpff_polar / 3.0, 1.0 / 3.0);
    b.inv_R_qq = 1.0 / sqrt(R_q * R_q + R_q * R_q);
    b.inv_R_pq = 1.0 / sqrt(R_p * R_p + R_q * R_q);
    b.inv_R_pp = 1.0 / sqrt(R_p * R_p + R_p * R_p);
    for (int d = 0; d < 3; ++d) b.efield[d] = p->efield[d];
}

static int validate(const capac_params* p) {
    if (!p || p->n_atoms <= 0) { capac_set_error("n_atoms must be positive"); return CAPAC_ERR_ARG; }
    if (p->n_metal < 0 || p->n_metal > p->n_atoms) {
        capac_set_error("n_metal %d outside [0, n_atoms]", p->n_metal);
        return CAPAC_ERR_ARG;
    }
    if (p->n_types < 1 || p->n_types > 8) {
        capac_set_error("%d vdW types; this build supports 1..8", p->n_types);
        return CAPAC_ERR_UNSUPPORTED;
    }
    for (int k = 0; k < p->n_types * p->n_types; ++k)
        if (p->nb_funct[k] < 1 || p->nb_funct[k] > 4) {
            capac_set_error("Error: unsupported vdW type: %d", p->nb_funct[k]);
            return CAPAC_ERR_UNSUPPORTED;
        }
    if (p->n_mols >= (1 << 26)) {
        // the neighbour build packs molecule * 32 + index into one int
        capac_set_error("%d molecules: at most 2^26 are supported", p->n_mols);
        return CAPAC_ERR_UNSUPPORTED;
    }
    for (int i = 0; i < p->n_atoms; ++i) {
        if (p->atomtype[i] < 0 || p->atomtype[i] >= p->n_types) {
            capac_set_error("atom %d has type %d outside the pair table", i, p->atomtype[i]);
            return CAPAC_ERR_ARG;
        }
        if (p->local_idx[i] < 0 || p->local_idx[i] >= 32) {
            capac_set_error("atom %d: molecules of more than 32 atoms are not supported", i);
            return CAPAC_ERR_UNSUPPORTED;
        }
    }
    for (int i = 0; i < p->n_metal; ++i) {
        int m = p->mol_id[i];
        if (p->excl_mask[i] != 0 && p->mol_last[m] > p->mol_first[m]) {
            capac_set_error("metal atom %d has intramolecular exclusions; metal atoms must not "
                            "exclude each other", i);
            return CAPAC_ERR_UNSUPPORTED;
        }
    }
    for (int k = 0; k < p->n_nps; ++k) {
        if (p->np_first[k] < 0 || p->np_last[k] >= p->n_metal || p->np_first[k] > p->np_last[k]) {
            capac_set_error("nanoparticle %d range %d..%d outside the metal atoms", k, p->np_first[k],
                            p->np_last[k]);
            return CAPAC_ERR_ARG;
        }
        for (int l = 0; l < k; ++l)
            if (p->np_first[k] <= p->np_last[l] && p->np_first[l] <= p->np_last[k]) {
                capac_set_error("nanoparticles %d and %d overlap", l, k);
                return CAPAC_ERR_ARG;
            }
    }
    {
        // bonded terms: each term's atoms inside its own molecule (the bonded
        // kernels give a molecule one thread and let it write its atoms)
        const int cnt[5] = {p->n_bonds, p->n_pairs14, p->n_angles, p->n_dihedrals, p->n_vsites};
        const int* atoms[5] = {p->bond_atoms, p->pair14_atoms, p->angle_atoms, p->dihedral_atoms,
                               p->vsite_atoms};
        const int width[5] = {2, 2, 3, 4, 4};
        const void* par[5] = {p->bond_par, p->pair14_par, p->angle_par, p->dihedral_par, p->vsite_par};
        const char* what[5] = {"bond", "1-4 pair", "angle", "dihedral", "virtual site"};
        const int total = cnt[0] + cnt[1] + cnt[2] + cnt[3] + cnt[4];
        for (int t = 0; t < 5; ++t)
            if (cnt[t] < 0 || (cnt[t] > 0 && (!atoms[t] || !par[t]))) {
                capac_set_error("%s list: count %d without arrays", what[t], cnt[t]);
                return CAPAC_ERR_ARG;
            }
        if (total > 0 && !p->mol_terms) { capac_set_error("bonded terms without mol_terms"); return CAPAC_ERR_ARG; }
        if ((p->n_bonds > 0 && !p->bond_funct) || (p->n_angles > 0 && !p->angle_funct)) {
            capac_set_error("bonded terms without their funct arrays");
            return CAPAC_ERR_ARG;
        }
        const int M1 = p->n_mols + 1;
        for (int t = 0; total > 0 && t < 5; ++t) {
            const int* rg = p->mol_terms + (size_t)t * M1;
            if (rg[0] != 0 || rg[p->n_mols] != cnt[t]) {
                capac_set_error("%s ranges do not cover the %d terms", what[t], cnt[t]);
                return CAPAC_ERR_ARG;
            }
            for (int m = 0; m < p->n_mols; ++m) {
                if (rg[m + 1] < rg[m]) { capac_set_error("%s ranges decrease at molecule %d", what[t], m); return CAPAC_ERR_ARG; }
                for (int q = rg[m]; q < rg[m + 1]; ++q)
                    for (int u = 0; u < width[t]; ++u) {
                        const int a = atoms[t][(size_t)width[t] * q + u];
                        if (a < p->mol_first[m] || a > p->mol_last[m]) {
                            capac_set_error("%s %d: atom %d outside its molecule %d", what[t], q, a, m);
                            return CAPAC_ERR_ARG;
                        }
                    }
            }
        }
        for (int q = 0; q < p->n_bonds; ++q)
            if (p->bond_funct[q] != 1 && p->bond_funct[q] != 3) {
                capac_set_error("Error: unsupported bond function type: %d", p->bond_funct[q]);
                return CAPAC_ERR_UNSUPPORTED;
            }
        for (int q = 0; q < p->n_angles; ++q)
            if (p->angle_funct[q] != 1 && p->angle_funct[q] != 5) {
                capac_set_error("Error: unsupported angle function type: %d", p->angle_funct[q]);
                return CAPAC_ERR_UNSUPPORTED;
            }
    }
    for (int d = 0; d < 3; ++d)
        if (!(p->box[d] > 0.0)) { capac_set_error("box must be positive"); return CAPAC_ERR_ARG; }
    if (!(p->rcut > 0.0)) { capac_set_error("cut-off must be positive"); return CAPAC_ERR_ARG; }
    if (p->n_metal > 0 && p->use_cpff && (!(p->cpff_polar > 0.0) || !(p->cpff_capac > 0.0))) {
        capac_set_error("CPIM polarizability and capacitance must be positive");
        return CAPAC_ERR_ARG;
    }
    return 0;
}

extern "C" int capac_create(const capac_params* p, int precision, int device, capac_ctx** out) {
    if (!out) { capac_set_error("out is NULL"); return CAPAC_ERR_ARG; }
    *out = NULL;
    int rc = validate(p);
    if (rc) return rc;
    if (precision != CAPAC_PREC_FP64 && precision != CAPAC_PREC_FP32_REFINED) {
        capac_set_error("unknown precision %d", precision);
        return CAPAC_ERR_ARG;
    }
    CHECK_HIP(hipSetDevice(device));
    capac_ctx* c = (capac_ctx*)calloc(1, sizeof(capac_ctx));
    if (!c) { capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
    c->device = device;
    c->precision = precision;
    c->need_energy = 1;
    const int N = p->n_atoms, Nm = p->n_metal;
    BoxConst& b = c->bc;
    b.N = N; b.Nm = Nm; b.n_np = Nm > 0 ? p->n_nps : 0; b.n_types = p->n_types; b.n_mols = p->n_mols;
    derive_constants(c, p);
    c->n_mat = (Nm > 0) ? 4 * Nm + b.n_np : 0;
    c->n_cstr = p->n_cstr;
    c->ndf = p->ndf;
    c->fix_metal = p->fix_metal;
    c->nvt = p->nvt;
    c->n_heating = p->n_heating;
    c->dt = p->dt;
    c->tau_t = p->tau_t;
    c->ref_temp0 = p->ref_temp;
    c->ext_temp = p->ref_temp;
    c->cur_ref_temp = p->ref_temp;
    // n_cstr and use_cpff bookkeeping
    int use_cpff = p->use_cpff && Nm > 0;
    if (!use_cpff) c->n_mat = 0;

    rc = 0;
#define TRY(x) do { if ((rc = (x)) != 0) { capac_destroy(c); return rc; } } while (0)
    TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess ? 0 : CAPAC_ERR_HIP);
    TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) == hipSuccess ? 0 : CAPAC_ERR_HIP);
    TRY(hipEventCreateWithFlags(&c->ev_rows, hipEventDisableTiming) == hipSuccess ? 0 : CAPAC_ERR_HIP);
    TRY(hipEventCreateWithFlags(&c->ev_side, hipEventDisableTiming) == hipSuccess ? 0 : CAPAC_ERR_HIP);
    size_t n3 = 3 * (size_t)N;
    TRY(upload(&c->r, (const double*)NULL, n3));
    TRY(upload(&c->v, (const double*)NULL, n3));
    TRY(upload(&c->f, (const double*)NULL, n3));
    TRY(upload(&c->old_r, (const double*)NULL, n3));
    TRY(upload(&c->charge, p->charge, N));
    TRY(upload(&c->mass, p->mass, N));
    {
        double* im = (double*)malloc(sizeof(double) * N);
        for (int i = 0; i < N; ++i) im[i] = p->mass[i] > 0.0 ? 1.0 / p->mass[i] : 0.0;
        rc = upload(&c->inv_mass, (const double*)im, N);
        free(im);
        if (rc) { capac_destroy(c); return rc; }
    }
    TRY(upload(&c->type, p->atomtype, N));
    TRY(upload(&c->mol_id, p->mol_id, N));
    TRY(upload(&c->local_idx, p->local_idx, N));
    TRY(upload(&c->excl, p->excl_mask, N));
    {
        int nt = p->n_types;
        PairParam* t = (PairParam*)calloc(nt * nt, sizeof(PairParam));
        for (int k = 0; k < nt * nt; ++k) {
            t[k].funct = p->nb_funct[k];
            t[k].c6 = p->nb_c6[k];
            t[k].c12 = p->nb_c12[k];
            t[k].width = p->nb_a[k];
            t[k].inv_width = p->nb_a[k] != 0.0 ? 1.0 / p->nb_a[k] : 0.0;
            t[k].a = p->nb_a[k];
            t[k].b = p->nb_b[k];
            const double inv_rc12 = 1.0 / pow(p->rcut, 12.0), inv_rc6 = 1.0 / pow(p->rcut, 6.0);
            t[k].lj_fc = 12.0 * t[k].c12 * inv_rc12 - 6.0 * t[k].c6 * inv_rc6;
            t[k].lj_fc_rc = t[k].lj_fc / p->rcut;
            t[k].lj_ec = t[k].c12 * inv_rc12 - t[k].c6 * inv_rc6;
        }
        rc = upload(&c->pair, (const PairParam*)t, nt * nt);
        free(t);
        if (rc) { capac_destroy(c); return rc; }
        // the water pair kernel: every pair of non-metal atom types funct 1,
        // damped-shifted-force Coulomb and shifted-force LJ, alpha rc < 6
        // (erfc from the erfcx table over the whole cut-off); with the metal
        // types' pairs funct 1 or 4 (erf-damped vdW, width > 0) it takes
        // every row (pair_water = 2)
        char* used = (char*)calloc(nt > 0 ? nt : 1, 1);
        for (int i = p->n_metal; i < N; ++i)
            if (p->atomtype[i] >= 0 && p->atomtype[i] < nt) used[p->atomtype[i]] = 1;
        int water = p->use_coulomb == 1 && p->use_vdw == 1 && p->dsf_alpha >= 0.0 &&
                    p->dsf_alpha * p->rcut < 5.99 && nt <= 8;
        for (int a = 0; a < nt && water; ++a)
            for (int b2 = 0; b2 < nt; ++b2)
                if (used[a] && used[b2] && p->nb_funct[a * nt + b2] != 1) water = 0;
        if (water) {
            for (int i = 0; i < p->n_metal; ++i)
                if (p->atomtype[i] >= 0 && p->atomtype[i] < nt) used[p->atomtype[i]] = 1;
            int all = 1;
            for (int a = 0; a < nt && all; ++a)
                for (int b2 = 0; b2 < nt; ++b2) {
                    const int k = a * nt + b2, fn = p->nb_funct[k];
                    if (used[a] && used[b2] && fn != 1 && !(fn == 4 && p->nb_a[k] > 0.0)) all = 0;
                }
            if (all) water = 2;
            else                        // the LJ bits below over the non-metal partners only
                for (int i = 0; i < p->n_metal; ++i)
                    if (p->atomtype[i] >= 0 && p->atomtype[i] < nt) used[p->atomtype[i]] = 0;
            for (int i = p->n_metal; i < N; ++i)
                if (p->atomtype[i] >= 0 && p->atomtype[i] < nt) used[p->atomtype[i]] = 1;
        }
        int lj_types = 0;
        for (int a = 0; a < nt && a < 31; ++a)
            for (int b2 = 0; b2 < nt; ++b2)
                if (used[b2] && p->nb_funct[a * nt + b2] == 1 &&
                    (p->nb_c6[a * nt + b2] != 0.0 || p->nb_c12[a * nt + b2] != 0.0))
                    lj_types |= 1 << a;
        c->pair_lj_types = lj_types;
        free(used);
        c->pair_water = water;
        // metal-metal pairs that the pair term would evaluate to exact zeros
        char* mused = (char*)calloc(nt > 0 ? nt : 1, 1);
        int inert = 1;
        for (int i = 0; i < p->n_metal; ++i) {
            if (p->charge[i] != 0.0) inert = 0;
            if (p->atomtype[i] >= 0 && p->atomtype[i] < nt) mused[p->atomtype[i]] = 1;
        }
        for (int a = 0; a < nt && inert; ++a)
            for (int b2 = 0; b2 < nt; ++b2)
                if (mused[a] && mused[b2] &&
                    (p->nb_funct[a * nt + b2] != 1 || p->nb_c6[a * nt + b2] != 0.0 || p->nb_c12[a * nt + b2] != 0.0))
                    inert = 0;
        free(mused);
        c->pair_metal_inert = inert;
        if (water) TRY(upload(&c->xyzq, (const double*)NULL, 4 * (size_t)N));
    }
    TRY(upload(&c->mol_first, p->mol_first, p->n_mols));
    TRY(upload(&c->mol_last, p->mol_last, p->n_mols));
    TRY(upload(&c->mol_cstr, p->mol_cstr, p->n_mols + 1));
    c->n_bonds = p->n_bonds; c->n_pairs14 = p->n_pairs14; c->n_angles = p->n_angles;
    c->n_dihedrals = p->n_dihedrals; c->n_vsites = p->n_vsites;
    if (p->n_bonds + p->n_pairs14 + p->n_angles + p->n_dihedrals + p->n_vsites > 0) {
        TRY(upload(&c->mol_terms, p->mol_terms, 5 * ((size_t)p->n_mols + 1)));
        TRY(upload(&c->bond_atoms, p->bond_atoms, 2 * (size_t)p->n_bonds));
        TRY(upload(&c->bond_funct, p->bond_funct, (size_t)p->n_bonds));
        TRY(upload(&c->bond_par, p->bond_par, 3 * (size_t)p->n_bonds));
        TRY(upload(&c->pair14_atoms, p->pair14_atoms, 2 * (size_t)p->n_pairs14));
        TRY(upload(&c->pair14_par, p->pair14_par, 6 * (size_t)p->n_pairs14));
        TRY(upload(&c->angle_atoms, p->angle_atoms, 3 * (size_t)p->n_angles));
        TRY(upload(&c->angle_funct, p->angle_funct, (size_t)p->n_angles));
        TRY(upload(&c->angle_par, p->angle_par, 4 * (size_t)p->n_angles));
        TRY(upload(&c->dihedral_atoms, p->dihedral_atoms, 4 * (size_t)p->n_dihedrals));
        TRY(upload(&c->dihedral_par, p->dihedral_par, 6 * (size_t)p->n_dihedrals));
        TRY(upload(&c->vsite_atoms, p->vsite_atoms, 4 * (size_t)p->n_vsites));
        TRY(upload(&c->vsite_par, p->vsite_par, 3 * (size_t)p->n_vsites));
    }
    TRY(upload(&c->cstr_i, p->cstr_i, p->n_cstr));
    TRY(upload(&c->cstr_j, p->cstr_j, p->n_cstr));
    TRY(upload(&c->cstr_r, p->cstr_r, p->n_cstr));
    TRY(upload(&c->np_first, p->np_first, b.n_np));
    TRY(upload(&c->np_last, p->np_last, b.n_np));
    TRY(upload(&c->np_charge, p->np_charge, b.n_np));
    {
        int* npo = (int*)malloc(sizeof(int) * (Nm > 0 ? Nm : 1));
        for (int i = 0; i < Nm; ++i) npo[i] = -1;
        for (int k = 0; k < b.n_np; ++k)
            for (int i = p->np_first[k]; i <= p->np_last[k]; ++i) npo[i] = k;
        rc = upload(&c->np_of_metal, (const int*)npo, Nm);
        free(npo);
        if (rc) { capac_destroy(c); return rc; }
    }
    // cell list and neighbour rows (list radius rc + skin)
    c->skin_req = CAPAC_DEFAULT_SKIN;
    c->solve_crit = CAPAC_DEFAULT_SOLVE_CRIT;
    c->pred_order = precision == CAPAC_PREC_FP32_REFINED ? CAPAC_DEFAULT_PREDICTOR_FP32 : 0;
    {
        // fitting parameters start at the deck's values, widths derived as the reference does
        const double pi = 3.14159265358979323846;
        c->theta[0] = p->cpff_polar;
        c->theta[1] = p->cpff_capac;
        c->theta[2] = sqrt(2.0 / pi) * p->cpff_capac;
        c->theta[3] = pow(sqrt(2.0 / pi) * p->cpff_polar / 3.0, 1.0 / 3.0);
        c->theta[4] = p->qsc_eps;
        c->theta[5] = p->qsc_c;
    }
    c->nb_dirty = 1;
    TRY(nb_set_radius(c, c->skin_req));
    TRY(alloc_cells(c));
    TRY(upload(&c->r_build, (const double*)NULL, 3 * (size_t)N));
    TRY(upload(&c->nb_ctl, (const unsigned long long*)NULL, 8));
    TRY(upload(&c->cell_of, (const int*)NULL, N));
    TRY(upload(&c->cell_rank, (const int*)NULL, N));
    TRY(upload(&c->sorted_r, (const double*)NULL, 3 * (size_t)N));
    TRY(upload(&c->sorted_mol, (const int*)NULL, N));
    TRY(upload(&c->sorted_loc, (const int*)NULL, N));
    TRY(upload(&c->cell_atoms, (const int*)NULL, N));
    TRY(upload(&c->cell_tmp, (const int*)NULL, N));
    TRY(upload(&c->n_front, (const int*)NULL, N));
    TRY(upload(&c->n_back, (const int*)NULL, N));
    TRY(upload(&c->tail_list, (const int*)NULL, (size_t)(N - p->n_metal) + 1));
    {
        // room for the densest liquid or metal around a sphere of radius rc + skin
        double rl = b.rc + c->skin;
        double vol = 4.0 / 3.0 * 3.14159265358979 * rl * rl * rl;
        int est = (int)(vol * 140.0) + 64;
        int stride = ((est + 63) / 64) * 64;
        if (stride > N) stride = ((N + 63) / 64) * 64;
        if (stride < 64) stride = 64;
        TRY(alloc_nbr(c, stride));
    }
    // reductions: one partial segment per producer
    {
        const int metal_blocks = (Nm + 3) / 4 + 1;
        c->seg_cap[SEG_PAIR] = pair_partial_blocks(N);
        c->seg_cap[SEG_SC] = metal_blocks;
        c->seg_cap[SEG_MW] = Nm + 1;        // one workgroup per metal atom
        // the metal-metal tiles, then one slot per block of 64 metal atoms
        c->seg_cap[SEG_MM] = (int)(tile_count(tile_blocks(Nm)) + tile_blocks(Nm));
        c->seg_cap[SEG_RATTLE] = (p->n_mols + 63) / 64 + 1;      // workgroups of 64 or more
        c->seg_cap[SEG_MOM] = (N + 255) / 256 + 1;
        c->seg_cap[SEG_BONDED] = bonded_blocks(p->n_mols) + 1;
        int total = 0;
        for (int s = 0; s < SEG_COUNT; ++s) { c->seg_off[s] = total; total += c->seg_cap[s]; }
        c->n_partial = total;
        c->n_pending = 0;
        TRY(upload(&c->partial, (const double*)NULL, (size_t)total * 16));
        // one buffer for everything the host reads back (layout at kRedWords)
        TRY(upload(&c->red, (const double*)NULL, kRedWords));
        TRY(hipMemset(c->red, 0, kRedWords * sizeof(double)) == hipSuccess ? 0 : CAPAC_ERR_HIP);
        TRY(hipHostMalloc(&c->red_host, kRedWords * sizeof(double), hipHostMallocDefault) == hipSuccess
                ? 0 : CAPAC_ERR_HIP);
        c->nb_overflow = (int*)(c->red + kRedFlags);
        c->md = c->red + kRedMd;
        c->md_status = (int*)(c->red + kRedStatus);
    }
    TRY(upload(&c->isd, (const double*)NULL, Nm));
    TRY(upload(&c->f_sc, (const double*)NULL, 3 * (size_t)Nm));
    TRY(upload(&c->mm_part, (const double*)NULL, (size_t)tile_blocks(Nm) * tile_blocks(Nm) * 3 * kTileEdge));
    {
        double qm = (double)p->ndf * (K_BOLTZ * p->ref_temp) * p->tau_t * p->tau_t;  // mpi-main.c:552
        double h[32] = {0};
        h[25] = qm;
        h[22] = 1.0;
        TRY(hipMemcpy(c->md, h, sizeof h, hipMemcpyHostToDevice) == hipSuccess ? 0 : CAPAC_ERR_HIP);
    }
    if (c->n_mat > 0) {
        TRY(upload(&c->b, (const double*)NULL, c->n_mat));
        TRY(upload(&c->x, (const double*)NULL, c->n_mat));
        TRY(upload(&c->x_backup, (const double*)NULL, c->n_mat));
        for (int k = 0; k < 3; ++k) TRY(upload(&c->x_hist[k], (const double*)NULL, c->n_mat));
        TRY(upload(&c->diag, (const double*)NULL, c->n_mat));
        // the FP64 path stores no relay (its product forms the blocks)
        c->relay_ld = ((Nm + 63) / 64) * 64;
        if (precision == CAPAC_PREC_FP32_REFINED) {
            c->bj_every = CAPAC_BJ_EVERY_DEFAULT;
            TRY(relay32_alloc(c, CAPAC_LAYOUT_AUTO));
            TRY(upload(&c->metal_u, (const double*)NULL, 4 * ((size_t)Nm + 1)));
        }
        TRY(solver_create(c));
        TRY(solver_noalloc_init(device));
    }
    for (int k = 0; k < 16; ++k) TRY(hipEventCreate(&c->ev[k]) == hipSuccess ? 0 : CAPAC_ERR_HIP);
#undef TRY
    *out = c;
    return 0;
}

extern "C" void capac_destroy(capac_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->r, c->v, c->f, c->old_r, c->charge, c->mass, c->inv_mass, c->type, c->mol_id,
                    c->local_idx, c->excl, c->pair, c->mol_first, c->mol_last, c->mol_cstr,
                    c->cstr_i, c->cstr_j, c->cstr_r, c->np_first, c->np_last, c->np_of_metal,
                    c->np_charge, c->cell_of, c->cell_count, c->cell_start, c->cell_atoms,
                    c->cell_tmp, c->cell_fill, c->cell_rank, c->sorted_r, c->sorted_mol,
                    c->sorted_loc, c->stencil, c->n_front, c->n_back, c->nbr, c->r_build, c->nb_ctl,
                    c->partial, c->red, c->isd, c->f_sc, c->mm_part, c->b, c->x, c->x_backup, c->diag, c->relay32, c->relay_part,
                    c->relay_cnt, c->metal_u, c->x_hist[0], c->x_hist[1], c->x_hist[2], c->mol_terms, c->bond_atoms,
                    c->bond_funct, c->bond_par, c->pair14_atoms, c->pair14_par, c->angle_atoms,
                    c->angle_funct, c->angle_par, c->dihedral_atoms, c->dihedral_par, c->vsite_atoms,
                    c->vsite_par, c->tail_list, c->xyzq};
    for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; ++k)
        if (ptrs[k]) (void)hipFree(ptrs[k]);
    bj_free(c);
    if (c->red_host) (void)hipHostFree(c->red_host);
    solver_destroy(c);
    fit_destroy(c);
    em_destroy(c);
    for (int k = 0; k < 16; ++k)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_rows) (void)hipEventDestroy(c->ev_rows);
    if (c->ev_side) (void)hipEventDestroy(c->ev_side);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    free(c);
}

// ------------------------------------------------------------------ state

extern "C" int capac_set_state(capac_ctx* c, const double* r, const double* v, const double* x) {
    if (!c) return CAPAC_ERR_ARG;
    size_t n3 = 3 * (size_t)c->bc.N;
    if (r) CHECK_HIP(hipMemcpyAsync(c->r, r, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (v) CHECK_HIP(hipMemcpyAsync(c->v, v, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (x && c->n_mat)
        CHECK_HIP(hipMemcpyAsync(c->x, x, c->n_mat * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    if (r) c->nb_dirty = 1;
    if (r || x) { c->pred_count = 0; c->pred_pending = 0; }   // the stored solutions no longer describe this state
    return 0;
}

extern "C" int capac_get_state(capac_ctx* c, double* r, double* v, double* x) {
    if (!c) return CAPAC_ERR_ARG;
    size_t n3 = 3 * (size_t)c->bc.N;
    if (r) CHECK_HIP(hipMemcpyAsync(r, c->r, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (v) CHECK_HIP(hipMemcpyAsync(v, c->v, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (x && c->n_mat)
        CHECK_HIP(hipMemcpyAsync(x, c->x, c->n_mat * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int capac_get_forces(capac_ctx* c, double* f) {
    if (!c || !f) return CAPAC_ERR_ARG;
    CHECK_HIP(hipMemcpyAsync(f, c->f, 3 * (size_t)c->bc.N * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

static int nb_regrid(capac_ctx* c);

extern "C" int capac_set_box(capac_ctx* c, const double box[3]) {
    if (!c || !box) return CAPAC_ERR_ARG;
    for (int d = 0; d < 3; ++d) {
        if (!(box[d] > 0.0)) { capac_set_error("box must be positive"); return CAPAC_ERR_ARG; }
        c->bc.box[d] = box[d];
        c->bc.half[d] = 0.5 * box[d];
    }
    c->pred_count = 0; c->pred_pending = 0;
    return nb_regrid(c);
}

// list radius and cell grid for the current box and requested skin
static int nb_regrid(capac_ctx* c) {
    int old = c->n_cell;
    int rc = nb_set_radius(c, c->skin_req);
    if (rc) return rc;
    c->nb_dirty = 1;
    if (c->n_cell > old) return alloc_cells(c);
    return 0;
}

extern "C" int capac_set_skin(capac_ctx* c, double skin) {
    if (!c) return CAPAC_ERR_ARG;
    if (!(skin >= 0.0) || !(skin < 0.5 * c->bc.rc + 1.0)) {
        capac_set_error("skin must be >= 0 and finite");
        return CAPAC_ERR_ARG;
    }
    c->skin_req = skin;
    return nb_regrid(c);
}

extern "C" double capac_skin(const capac_ctx* c) { return c ? c->skin : -1.0; }

extern "C" int capac_set_theta(capac_ctx* c, const double theta[CAPAC_NTHETA]) {
    if (!c || !theta) return CAPAC_ERR_ARG;
    for (int k = 0; k < CAPAC_NTHETA; ++k)
        if (!(theta[k] > 0.0) || !(theta[k] < 1e300)) {
            capac_set_error("theta[%d] = %g must be positive and finite", k, theta[k]);
            return CAPAC_ERR_ARG;
        }
    memcpy(c->theta, theta, sizeof(double) * CAPAC_NTHETA);
    fit_apply_theta(c);
    return 0;
}

extern "C" int capac_get_theta(capac_ctx* c, double theta[CAPAC_NTHETA]) {
    if (!c || !theta) return CAPAC_ERR_ARG;
    memcpy(theta, c->theta, sizeof(double) * CAPAC_NTHETA);
    return 0;
}

extern "C" int capac_fit_frame(capac_ctx* c, const double* f_ref, double e_ref, double w_e, double* loss,
                               double grad[CAPAC_NTHETA]) {
    if (!c || !f_ref) return CAPAC_ERR_ARG;
    return fit_frame_loss_grad(c, f_ref, e_ref, w_e, loss, grad);
}

extern "C" int capac_set_fp32_layout(capac_ctx* c, int layout) {
    if (!c) return CAPAC_ERR_ARG;
    if (c->precision != CAPAC_PREC_FP32_REFINED || !c->n_mat) {
        capac_set_error("the relay layout applies to the FP32-refined path with metal atoms only");
        return CAPAC_ERR_ARG;
    }
    if (layout != CAPAC_LAYOUT_AUTO && layout != CAPAC_LAYOUT_ROWS && layout != CAPAC_LAYOUT_TILES) {
        capac_set_error("unknown relay layout %d", layout);
        return CAPAC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return relay32_alloc(c, layout);
}

extern "C" int capac_fp32_layout(const capac_ctx* c) {
    return c && c->precision == CAPAC_PREC_FP32_REFINED && c->n_mat ? c->relay_layout : -1;
}

extern "C" int capac_set_block_precond(capac_ctx* c, int every) {
    if (!c) return CAPAC_ERR_ARG;
    if (c->precision != CAPAC_PREC_FP32_REFINED || !c->n_mat) {
        capac_set_error("the block preconditioner applies to the FP32-refined path with metal atoms only");
        return CAPAC_ERR_ARG;
    }
    if (every < 0) {
        capac_set_error("block preconditioner interval must be >= 0 (0: Jacobi), got %d", every);
        return CAPAC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    CHECK_HIP(hipStreamSynchronize(c->stream));
    c->bj_every = every;
    return bj_alloc(c);
}

extern "C" int capac_block_precond(const capac_ctx* c) {
    return c && c->precision == CAPAC_PREC_FP32_REFINED && c->n_mat ? c->bj_every : -1;
}

extern "C" int capac_block_precond_inverse(capac_ctx* c, int block, float* out) {
    if (!c || !out) return CAPAC_ERR_ARG;
    if (c->precision != CAPAC_PREC_FP32_REFINED || !c->n_mat) {
        capac_set_error("the block preconditioner applies to the FP32-refined path with metal atoms only");
        return CAPAC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    return bj_copy_block(c, block, out);
}

extern "C" int capac_block_precond_coarse(capac_ctx* c, int* n, double* ainv, float* basis) {
    if (!c || !n) return CAPAC_ERR_ARG;
    if (c->precision != CAPAC_PREC_FP32_REFINED || !c->n_mat) {
        capac_set_error("the block preconditioner applies to the FP32-refined path with metal atoms only");
        return CAPAC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    return bj_copy_coarse(c, n, ainv, basis);
}

extern "C" int capac_set_predictor(capac_ctx* c, int order) {
    if (!c) return CAPAC_ERR_ARG;
    if (order < 0 || order > 2) {
        capac_set_error("predictor order must be 0, 1 or 2");
        return CAPAC_ERR_ARG;
    }
    c->pred_order = order;
    return 0;
}

extern "C" int capac_predictor(const capac_ctx* c) { return c ? c->pred_order : -1; }

extern "C" int capac_set_solve_tolerance(capac_ctx* c, double rtr_over_btb) {
    if (!c) return CAPAC_ERR_ARG;
    if (!(rtr_over_btb > 0.0) || !(rtr_over_btb <= 1.0)) {
        capac_set_error("solve tolerance must be in (0, 1]");
        return CAPAC_ERR_ARG;
    }
    c->solve_crit = rtr_over_btb;
    return 0;
}

// ------------------------------------------------------------------ force

static int fetch_red(capac_ctx* c) {
    int rc = flush_reductions(c);
    if (rc) return rc;
    // the reduction outputs and the overflow flags in one transfer
    CHECK_HIP(hipMemcpyAsync(c->red_host, c->red, kRedMd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

static int nb_overflowed(capac_ctx* c) { return ((int*)(c->red_host + kRedFlags))[0] != 0; }

static int grow_nbr(capac_ctx* c) {
    int stride = c->nb_stride * 2;
    if (stride > ((c->bc.N + 63) / 64) * 64 + 64) stride = ((c->bc.N + 63) / 64) * 64 + 64;
    return alloc_nbr(c, stride);
}

static void mark(capac_ctx* c, int k) {
    if (c->timing_on) (void)hipEventRecord(c->ev[k], c->stream);
}

// grid of the pair kernel while it runs beside the solve.  Below 1024 metal
// atoms the solve is a chain of short, latency-bound launches: the pair
// kernel gets one or one and a half workgroups per compute unit, so that the
// solver's workgroups always find free slots (fuller grids delay each of
// them), the larger share when the warm-start predictor has made the solve
// short.  Measured on MI355X (ms per MD step): Au561 FP32-refined with the
// quadratic predictor 0.877 at 256 workgroups, 0.827 at 384, 0.965 at 512,
// 0.925 at 768; FP64 1.589 at 256, 1.603 at 384, 1.758 at 512.  Once the
// solve had shortened further (0.33 ms per step), 384 workgroups left the pair kernel ~12 us
// behind it; 448 let it finish in time: 0.562-0.571 against 0.580-0.581 ms
// per Au561 step on one box (320: 0.584-0.592).
// From 1024 metal atoms the pair kernel starts after the assembly (see
// force_compute) with two workgroups per compute unit: it then shares every
// compute unit with the solve's launches instead of holding their slots.
// Au2057 FP32-refined with the quadratic predictor, ms per MD step on one
// box: 2.53 at 256 workgroups, 2.22 at 384, 2.01 at 512, 2.10 at 640, 2.23
// at 768, 2.15 at 1024, 2.23 at 1536, 2.20 with the full grid (2048).  With
// the water pair kernel (two launches, the first at most 132 VGPRs) and the
// three-launch GMRES column: 1.300 at 384, 1.236 at 416, 1.242 at 432,
// 1.199-1.205 at 448, 1.214 at 456, 1.256 at 464, 1.288 at 480, 1.223-1.227
// at 512, 1.317 at 640 (the pair launches then end 0.03-0.05 ms after the
// solve at 448, and the solve takes 0.75 ms per step against 0.81 at 512).
#ifndef CAPAC_SIDE_BLOCKS_LARGE
#define CAPAC_SIDE_BLOCKS_LARGE 448
#endif
static int side_pair_blocks(const capac_ctx* c) {
    if (c->bc.Nm >= 1024) return CAPAC_SIDE_BLOCKS_LARGE;
    return c->pred_order > 0 ? 448 : 256;
}

// pair forces (and Sutton-Chen without CPIM unknowns) on the side stream,
// after the rows are built (event ev_rows); they write f first, so everything
// that accumulates into f afterwards waits for ev_side.  Their reductions are
// queued here and launched on the main stream once it has waited.
static int side_forces(capac_ctx* c) {
    int rc = flush_reductions(c);
    if (rc) return rc;
    CHECK_HIP(hipEventRecord(c->ev_rows, c->stream));
    CHECK_HIP(hipStreamWaitEvent(c->side, c->ev_rows, 0));
    hipStream_t main_stream = c->stream;
    c->stream = c->side;
    mark(c, 11);                     // events 11-12: the side stream's pair phase
    if (has_bonded(c)) {
        // the reference computes the bonded forces first (force.c:2163-2168);
        // the pair forces then add to them
        rc = fill_d(c, c->f, 3 * (size_t)c->bc.N, 0.0);
        if (!rc) rc = bonded_forces(c);
        if (!rc) rc = pair_nonbonded(c, c->f, side_pair_blocks(c), 1);
    } else {
        rc = pair_nonbonded(c, NULL, side_pair_blocks(c), 1);
    }
    mark(c, 12);
    // (with CPIM unknowns the Sutton-Chen force runs on the main stream)
    if (!rc && c->bc.Nm > 0 && !c->n_mat) rc = sc_forces(c);
    c->stream = main_stream;
    if (rc) return rc;
    CHECK_HIP(hipEventRecord(c->ev_side, c->side));
    return 0;
}

static void take_force_results(capac_ctx* c);

// one mpi_force: forces into c->f, pot/vir into c->pot/c->vir.
// deferred (MD steps with CPIM unknowns): the call does not synchronise.  The
// neighbour-row flags come back with the solver's last synchronisation (they
// share its transfer), and the caller collects the energies with
// take_force_results after its own readback at the end of the step.
static int force_compute(capac_ctx* c, bool deferred = false) {
    int rc;
    deferred = deferred && c->n_mat > 0;
    c->timing.n_force_calls++;          // every call, also those whose results stay on the device
    for (int attempt = 0; attempt < 8; ++attempt) {
        mark(c, 0);
        // virtual sites from their atoms first (force.c:2128-2134)
        if ((rc = vsite_build(c))) return rc;
        if ((rc = flush_reductions(c))) return rc;     // nothing queued may land after the clear
        if ((rc = force_prologue(c, 256))) return rc;
        if ((rc = c->nb_dirty ? nb_build(c) : nb_build_if_needed(c))) return rc;
        c->nb_dirty = 0;
        mark(c, 1);
        // A pair launch over the large decks' grid takes every compute unit
        // for its whole duration, and launches queued behind it on the main
        // stream wait for free slots: at Au2057 the 14 us RHS launch measured
        // 450 us on average.  So there the pair kernel starts once the RHS
        // and the assembly are done and overlaps only the solve; the capped
        // grid of the small decks leaves slots free and starts at once.
        const bool side_after_matrix = c->n_mat && c->bc.Nm >= 1024;
        // With CPIM unknowns the Sutton-Chen force goes to f_sc on the main
        // stream, right after the side stream has started (it needs only the
        // rows), and the metal-metal CPIM launch adds it: the side stream,
        // which ends after the solve at Au2057, loses two launches.
        auto sc_main = [&]() -> int {
            if (!c->n_mat) return 0;
            mark(c, 2);
            if (sc_in_force_launches(c)) {
                // (the passes ride in the RHS/relay and CPIM force launches;
                // the SC span stays recorded, empty)
                mark(c, 3);
                return 0;
            }
            int r2 = sc_forces(c, c->f_sc);
            mark(c, 3);
            return r2;
        };
        if (!side_after_matrix && ((rc = side_forces(c)) || (rc = sc_main()))) return rc;
        int solve_rc = 0;
        if (c->n_mat) {
            // (one launch for both on the FP32 path's TILES layout: the RHS
            // span then covers the relay too)
            if ((rc = cpim_rhs_assemble(c))) return rc;
            mark(c, 4);
            mark(c, 5);
            if (side_after_matrix && ((rc = side_forces(c)) || (rc = sc_main()))) return rc;
            solve_rc = solver_run(c, c->b, c->x);
            if (solve_rc && solve_rc != CAPAC_ERR_SOLVER) {
                (void)hipStreamSynchronize(c->side);
                return solve_rc;
            }
        } else {
            mark(c, 4); mark(c, 5);
        }
        if (deferred && nb_overflowed(c)) {
            // a neighbour row overflowed (rare): let the side stream's readers
            // of the rows finish, then grow, restore the warm start and redo
            CHECK_HIP(hipStreamSynchronize(c->side));
            c->nb_dirty = 1;
            if ((rc = grow_nbr(c))) return rc;
            if ((rc = copy_d(c, c->x_backup, c->x, c->n_mat))) return rc;
            if (attempt == 7) { capac_set_error("neighbour rows keep overflowing"); return CAPAC_ERR_HIP; }
            continue;
        }
        mark(c, 6);
        CHECK_HIP(hipStreamWaitEvent(c->stream, c->ev_side, 0));
        if (c->n_mat && (rc = cpim_forces(c))) return rc;
        rc = solve_rc;
        // (the external field was added with the pair forces)
        // forces on virtual sites go to their atoms once everything is in
        // (force.c:2300-2306)
        int rc2 = vsite_spread(c, c->f);
        if (rc2) return rc2;
        mark(c, 7);
        if (deferred) break;
        rc2 = fetch_red(c);
        if (rc2) return rc2;
        if (!nb_overflowed(c)) break;
        // a neighbour row overflowed: grow, restore the warm start, redo
        c->nb_dirty = 1;
        if ((rc2 = grow_nbr(c))) return rc2;
        if (c->n_mat && (rc2 = copy_d(c, c->x_backup, c->x, c->n_mat))) return rc2;
        if (attempt == 7) { capac_set_error("neighbour rows keep overflowing"); return CAPAC_ERR_HIP; }
    }
    if (!deferred) take_force_results(c);
    return rc;
}

// energies, virial and phase timings of the last force call, from the
// readback buffer (after a synchronisation that covers the call)
static void take_force_results(capac_ctx* c) {
    const double* h = c->red_host;
    memset(c->pot, 0, sizeof c->pot);
    memset(c->vir, 0, sizeof c->vir);
    c->pot[7] = h[0];
    c->pot[8] = h[1];
    c->pot[1] = h[16];
    c->pot[10] = h[32];
    c->pot[11] = h[33];
    c->pot[9] = h[48] + h[11];
    c->pot[12] = h[49];
    c->pot[13] = h[50];
    c->pot[14] = h[51];
    for (int k = 0; k < 9; ++k) c->vir[k] = h[2 + k] + h[17 + k] + h[34 + k] + h[52 + k];
    if (has_bonded(c)) {
        c->pot[2] = h[112];
        c->pot[3] = h[113];
        c->pot[4] = h[114];
        c->pot[7] += h[115];
        c->pot[8] += h[116];
        for (int k = 0; k < 9; ++k) c->vir[k] += h[117 + k];
    }
    if (c->timing_on) {
        float ms;
        // (from, to, category): build, RHS, matrix, solve, CPIM force and SC
        // (with CPIM unknowns) on the main stream; pair on the side stream,
        // overlapping the solve
        const int spans[7][3] = {{0, 1, 10}, {1, 4, 6}, {4, 5, 5}, {5, 6, 7}, {6, 7, 8}, {11, 12, 4}, {2, 3, 2}};
        for (int k = 0; k < 7; ++k) {
            if (hipEventElapsedTime(&ms, c->ev[spans[k][0]], c->ev[spans[k][1]]) == hipSuccess)
                c->timing.t[spans[k][2]] += 1e-3 * ms;
            else
                (void)hipGetLastError();    // a span this call did not record: no time, and no error left behind
        }
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[7]) == hipSuccess) c->timing.t[0] += 1e-3 * ms;
    }
}

extern "C" int capac_force(capac_ctx* c, double pot[CAPAC_NPOT], double vir[9], double* f) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = force_compute(c);
    if (rc && rc != CAPAC_ERR_SOLVER) return rc;
    if (pot) memcpy(pot, c->pot, sizeof c->pot);
    if (vir) memcpy(vir, c->vir, sizeof c->vir);
    if (f) {
        int rc2 = capac_get_forces(c, f);
        if (rc2) return rc2;
    }
    return rc;
}

// ------------------------------------------------------------------ sub-operators

static int build_until_fits(capac_ctx* c) {
    for (int attempt = 0; attempt < 8; ++attempt) {
        int rc = nb_build(c);
        if (rc) return rc;
        if ((rc = fetch_red(c))) return rc;
        c->nb_dirty = 0;
        if (!nb_overflowed(c)) return 0;
        c->nb_dirty = 1;
        if ((rc = grow_nbr(c))) return rc;
    }
    capac_set_error("neighbour rows keep overflowing");
    return CAPAC_ERR_HIP;
}

extern "C" int capac_neighbours(capac_ctx* c, int* count, int* list, int max_per_atom) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = build_until_fits(c);
    if (rc) return rc;
    const int N = c->bc.N, S = c->nb_stride;
    int* nf = (int*)malloc(sizeof(int) * N);
    int* nb = (int*)malloc(sizeof(int) * N);
    int* rows = (int*)malloc(sizeof(int) * (size_t)N * S);
    if (!nf || !nb || !rows) { free(nf); free(nb); free(rows); capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
    hipError_t e1 = hipMemcpy(nf, c->n_front, sizeof(int) * N, hipMemcpyDeviceToHost);
    hipError_t e2 = hipMemcpy(nb, c->n_back, sizeof(int) * N, hipMemcpyDeviceToHost);
    hipError_t e3 = hipMemcpy(rows, c->nbr, sizeof(int) * (size_t)N * S, hipMemcpyDeviceToHost);
    int mx = 0;
    if (e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess) {
        for (int i = 0; i < N; ++i) {
            int n = nf[i] + nb[i];
            if (n > mx) mx = n;
            if (count) count[i] = n;
            if (list && n <= max_per_atom) {
                int* out = list + (size_t)i * max_per_atom;
                memcpy(out, rows + (size_t)i * S, sizeof(int) * nf[i]);
                memcpy(out + nf[i], rows + (size_t)i * S + S - nb[i], sizeof(int) * nb[i]);
            }
        }
    } else {
        capac_set_error("neighbour download failed");
        mx = CAPAC_ERR_HIP;
    }
    free(nf); free(nb); free(rows);
    if (mx > max_per_atom && list) {
        capac_set_error("a neighbour row has %d entries, more than max_per_atom %d", mx, max_per_atom);
        return CAPAC_ERR_ARG;
    }
    return mx;
}

static void pot_out(capac_ctx* c, double* pot, double* vir, const int* pk, const int* hk, int npk,
                    int vir_at) {
    if (pot) {
        memset(pot, 0, sizeof(double) * CAPAC_NPOT);
        for (int k = 0; k < npk; ++k) pot[pk[k]] = c->red_host[hk[k]];
    }
    if (vir)
        for (int k = 0; k < 9; ++k) vir[k] = c->red_host[vir_at + k];
}

extern "C" int capac_nonbonded(capac_ctx* c, double* f, double pot[CAPAC_NPOT], double vir[9]) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = build_until_fits(c);
    if (!rc) rc = clear_red(c, 0, 256);
    if (!rc) rc = pair_nonbonded(c, NULL);
    if (!rc) rc = fetch_red(c);
    if (rc) return rc;
    const int pk[2] = {7, 8}, hk[2] = {0, 1};
    pot_out(c, pot, vir, pk, hk, 2, 2);
    return f ? capac_get_forces(c, f) : 0;
}

extern "C" int capac_bonded(capac_ctx* c, double* f, double pot[CAPAC_NPOT], double vir[9]) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = zero_outputs(c);
    if (!rc) rc = bonded_forces(c);
    if (!rc) rc = fetch_red(c);
    if (rc) return rc;
    const int pk[5] = {2, 3, 4, 7, 8}, hk[5] = {112, 113, 114, 115, 116};
    pot_out(c, pot, vir, pk, hk, 5, 117);
    return f ? capac_get_forces(c, f) : 0;
}

extern "C" int capac_vsite_build(capac_ctx* c, double* r) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = vsite_build(c);
    if (rc) return rc;
    c->nb_dirty = 1;                 // positions changed outside the MD path
    CHECK_HIP(hipStreamSynchronize(c->stream));
    if (r) CHECK_HIP(hipMemcpy(r, c->r, sizeof(double) * 3 * c->bc.N, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int capac_vsite_spread(capac_ctx* c, const double* f_in, double* f_out) {
    if (!c || !f_in || !f_out) return CAPAC_ERR_ARG;
    const size_t n3 = 3 * (size_t)c->bc.N;
    CHECK_HIP(hipMemcpyAsync(c->f, f_in, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
    int rc = vsite_spread(c, c->f);
    if (rc) return rc;
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return capac_get_forces(c, f_out);
}

extern "C" int capac_sutton_chen(capac_ctx* c, double* f, double pot[CAPAC_NPOT], double vir[9],
                                 double* isd) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = build_until_fits(c);
    if (!rc) rc = zero_outputs(c);
    if (!rc) rc = sc_forces(c);
    if (!rc) rc = fetch_red(c);
    if (rc) return rc;
    const int pk[1] = {1}, hk[1] = {16};
    pot_out(c, pot, vir, pk, hk, 1, 17);
    if (isd && c->bc.Nm)
        CHECK_HIP(hipMemcpy(isd, c->isd, sizeof(double) * c->bc.Nm, hipMemcpyDeviceToHost));
    return f ? capac_get_forces(c, f) : 0;
}

extern "C" int capac_cpim_rhs(capac_ctx* c, double* b) {
    if (!c) return CAPAC_ERR_ARG;
    if (!c->n_mat) { capac_set_error("no CPIM unknowns (no metal or use_cpff 0)"); return CAPAC_ERR_ARG; }
    int rc = build_until_fits(c);
    if (!rc) rc = cpim_rhs(c);
    if (rc) return rc;
    if (b) CHECK_HIP(hipMemcpyAsync(b, c->b, sizeof(double) * c->n_mat, hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int capac_cpim_assemble(capac_ctx* c, double* dense, double* diag) {
    if (!c) return CAPAC_ERR_ARG;
    if (!c->n_mat) { capac_set_error("no CPIM unknowns (no metal or use_cpff 0)"); return CAPAC_ERR_ARG; }
    int rc = cpim_assemble(c);
    if (rc) return rc;
    if (dense) {
        size_t n4 = 4 * (size_t)c->bc.Nm;
        double* d = NULL;
        CHECK_HIP(hipMalloc(&d, n4 * n4 * sizeof(double)));
        rc = cpim_dense_expand(c, d);
        if (!rc && hipMemcpyAsync(dense, d, n4 * n4 * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = CAPAC_ERR_HIP;
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(d);
        if (rc) return rc;
    }
    if (diag) CHECK_HIP(hipMemcpyAsync(diag, c->diag, sizeof(double) * c->n_mat, hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int capac_cpim_solve(capac_ctx* c, const double* b, double* x, int* iters, double* rTr,
                                double* bTb, int* kind) {
    if (!c) return CAPAC_ERR_ARG;
    if (!c->n_mat) { capac_set_error("no CPIM unknowns (no metal or use_cpff 0)"); return CAPAC_ERR_ARG; }
    size_t bytes = sizeof(double) * c->n_mat;
    if (b) CHECK_HIP(hipMemcpyAsync(c->b, b, bytes, hipMemcpyHostToDevice, c->stream));
    if (x) CHECK_HIP(hipMemcpyAsync(c->x, x, bytes, hipMemcpyHostToDevice, c->stream));
    int rc = cpim_assemble(c);
    if (rc) return rc;
    rc = solver_run(c, c->b, c->x);
    if (rc && rc != CAPAC_ERR_SOLVER) return rc;
    if (x) CHECK_HIP(hipMemcpyAsync(x, c->x, bytes, hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    if (iters) *iters = c->solve_iters;
    if (rTr) *rTr = c->solve_rTr;
    if (bTb) *bTb = c->solve_bTb;
    if (kind) *kind = c->solve_kind;
    return rc;
}

// capac_cpim_solve with the solver's work space in the library's
// file-scope device array instead of the context's allocations
extern "C" int capac_cpim_solve_noalloc(capac_ctx* c, const double* b, double* x, int* iters, double* rTr,
                                        double* bTb, int* kind) {
    if (!c) return CAPAC_ERR_ARG;
    if (!c->n_mat) { capac_set_error("no CPIM unknowns (no metal or use_cpff 0)"); return CAPAC_ERR_ARG; }
    size_t bytes = sizeof(double) * c->n_mat;
    if (b) CHECK_HIP(hipMemcpyAsync(c->b, b, bytes, hipMemcpyHostToDevice, c->stream));
    if (x) CHECK_HIP(hipMemcpyAsync(c->x, x, bytes, hipMemcpyHostToDevice, c->stream));
    int rc = cpim_assemble(c);
    if (rc) return rc;
    rc = solver_run_noalloc(c, c->b, c->x);
    if (rc && rc != CAPAC_ERR_SOLVER) return rc;
    if (x) CHECK_HIP(hipMemcpyAsync(x, c->x, bytes, hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    if (iters) *iters = c->solve_iters;
    if (rTr) *rTr = c->solve_rTr;
    if (bTb) *bTb = c->solve_bTb;
    if (kind) *kind = c->solve_kind;
    return rc;
}

extern "C" int capac_cpim_force(capac_ctx* c, const double* x, double* f, double pot[CAPAC_NPOT],
                                double vir[9]) {
    if (!c) return CAPAC_ERR_ARG;
    if (!c->n_mat) { capac_set_error("no CPIM unknowns (no metal or use_cpff 0)"); return CAPAC_ERR_ARG; }
    int rc = build_until_fits(c);
    if (!rc) rc = zero_outputs(c);
    if (rc) return rc;
    const double* xd = c->x;
    if (x) {
        CHECK_HIP(hipMemcpyAsync(c->x_backup, x, sizeof(double) * c->n_mat,
                                 hipMemcpyHostToDevice, c->stream));
        xd = c->x_backup;
    }
    rc = cpim_forces_x(c, xd);
    if (!rc) rc = fetch_red(c);
    if (rc) return rc;
    if (pot) {
        memset(pot, 0, sizeof(double) * CAPAC_NPOT);
        pot[9] = c->red_host[48];
        pot[10] = c->red_host[32];
        pot[11] = c->red_host[33];
        pot[12] = c->red_host[49];
        pot[13] = c->red_host[50];
        pot[14] = c->red_host[51];
    }
    if (vir)
        for (int k = 0; k < 9; ++k) vir[k] = c->red_host[34 + k] + c->red_host[52 + k];
    return f ? capac_get_forces(c, f) : 0;
}

extern "C" int capac_solve_stats(capac_ctx* c, int* iters, double* rTr, double* bTb, int* kind) {
    if (!c) return CAPAC_ERR_ARG;
    if (iters) *iters = c->solve_iters;
    if (rTr) *rTr = c->solve_rTr;
    if (bTb) *bTb = c->solve_bTb;
    if (kind) *kind = c->solve_kind;
    return 0;
}

// ------------------------------------------------------------------ MD

static int fetch_md(capac_ctx* c) {
    int rc = flush_reductions(c);
    if (rc) return rc;
    // the MD scalars, RATTLE flags and the RATTLE_2 virial (red[64..73)),
    // with everything between them, in one transfer
    CHECK_HIP(hipMemcpyAsync(c->red_host, c->red, kRedWords * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ------------------------------------------------------------------ minimisation

extern "C" int capac_em_begin(capac_ctx* c) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = em_begin(c);
    if (!rc) rc = fill_i(c, c->md_status, 4, 0);
    if (!rc) rc = fetch_md(c);
    return rc;
}

extern "C" int capac_em_step(capac_ctx* c, double length, double pot[CAPAC_NPOT], double* fmax, double* frms) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = em_move(c, length);
    if (!rc) rc = md_rattle1(c);
    if (!rc) rc = fill_d(c, c->v, 3 * (size_t)c->bc.N, 0.0);       // mpi-main.c:753-759
    if (rc) return rc;
    int frc = force_compute(c);
    if (frc && frc != CAPAC_ERR_SOLVER) return frc;
    rc = md_fmax(c);
    if (!rc) rc = fetch_md(c);
    if (rc) return rc;
    const int* st = (const int*)(c->red_host + kRedStatus);
    if (st[0]) { capac_set_error("RATTLE_1 not converged!"); return CAPAC_ERR_RATTLE1; }
    if (pot) {
        memcpy(pot, c->pot, sizeof c->pot);
        double e = 0.0;
        for (int k = 1; k < CAPAC_NPOT; ++k) e += pot[k];     // sum_potential (force.c:322-331)
        pot[0] = e;
    }
    if (fmax) *fmax = c->red_host[kRedMd + 23];
    if (frms) *frms = c->red_host[kRedMd + 24];
    return frc;
}

extern "C" int capac_em_direction(capac_ctx* c, int cg) {
    if (!c) return CAPAC_ERR_ARG;
    return em_direction(c, cg);
}

extern "C" int capac_md_begin(capac_ctx* c) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = md_remove_com_kinetic(c);
    if (!rc) rc = md_fmax(c);
    if (!rc) rc = fill_i(c, c->md_status, 4, 0);
    if (!rc) rc = clear_red(c, 64, 16);
    if (!rc) rc = fetch_md(c);
    if (rc) return rc;
    c->first_temp = c->red_host[kRedMd + 1];
    c->ext_temp = c->ref_temp0;
    c->cur_ref_temp = c->ref_temp0;
    // the first force call's solution starts the predictor's history
    c->pred_count = 0; c->pred_pending = 0;
    solver_push_history(c);
    return 0;
}

extern "C" int capac_md_run(capac_ctx* c, int first_step, int n_steps, int n_save, int stop_at_save) {
    if (!c) return CAPAC_ERR_ARG;
    int done = 0;
    // every other entry point forms the energies
    struct EnergyReset {
        capac_ctx* c;
        ~EnergyReset() { c->need_energy = 1; }
    } energy_reset{c};
    for (int step = first_step; step < first_step + n_steps; ++step) {
        double ref_temp = step < c->n_heating
                              ? c->first_temp + (c->ext_temp - c->first_temp) * step / c->n_heating
                              : c->ext_temp;
        c->cur_ref_temp = ref_temp;
        double kT = K_BOLTZ * ref_temp;
        int rc = 0;
        mark(c, 8);
        rc = md_first_half(c, kT);
        if (!rc) rc = md_rattle1(c);
        if (rc) return rc;
        solver_predict(c);
        // A step that ends without a synchronisation (below) has its
        // energies and virial read by nobody: the reference prints and writes
        // them at save steps only (mpi-main.c:987-1000), and the integrator
        // uses forces and kinetic energies alone.  Its pair launch forms
        // forces only.
        const bool save = n_save > 0 && step % n_save == 0;
        const bool last = step + 1 == first_step + n_steps;
        c->need_energy = (save || last || c->timing_on || !c->n_mat) ? 1 : 0;
        c->hist_dst = c->n_mat ? c->x_hist[2] : NULL;   // solver_push_history's next newest buffer
        c->hist_copied = NULL;
        rc = force_compute(c, true);
        c->hist_dst = NULL;
        if (c->n_mat) {
            // the solve's synchronisation brought the RATTLE flags back (this
            // step's RATTLE_1 and the previous step's RATTLE_2; they stay set
            // until capac_md_begin).  They are checked before an error of the
            // force call is returned: the reference stops at the RATTLE
            // failure, which comes first (force.c:1990-1994, 2083-2087).  A
            // call that failed before its solve leaves the flags of an
            // earlier synchronisation, which were clear or already reported.
            const int* st = (const int*)(c->red_host + kRedStatus);
            if (st[0]) { capac_set_error("RATTLE_1 not converged!"); return CAPAC_ERR_RATTLE1; }
            if (st[1]) { capac_set_error("RATTLE_2 not converged!"); return CAPAC_ERR_RATTLE2; }
        }
        if (rc) return rc;
        solver_push_history(c);
        mark(c, 9);
        if (!rc) rc = md_rattle2(c, true);     // the half kick, then RATTLE_2

        // (an unobserved step runs the next step's first-half chain in its
        // own chain launch: the next step's temperature from the same ramp)
        double kT_next = 0.0;
        if (!c->need_energy) {
            const int s1 = step + 1;
            kT_next = K_BOLTZ * (s1 < c->n_heating
                                     ? c->first_temp + (c->ext_temp - c->first_temp) * s1 / c->n_heating
                                     : c->ext_temp);
        }
        if (!rc) rc = md_second_half(c, kT, kT_next);
        if (!rc && save) rc = md_wrap(c);
        if (!rc && save) rc = md_fmax(c);
        mark(c, 10);
        if (rc) return rc;
        // A step whose results nobody reads ends without a synchronisation,
        // so the next step's first launches queue behind this step's last
        // ones: not a save step, not the last step of this call, no HIP-event
        // timing (read per step), and a force call that synchronised in its
        // solve, which also reports the RATTLE flags one step later at most.
        // The potential energies of such a step are not brought back.
        if (!c->need_energy) {
            ++done;
            continue;
        }
        rc = fetch_md(c);
        if (rc) return rc;
        if (c->n_mat) take_force_results(c);
        if (c->timing_on) {
            float a, b2;
            if (hipEventElapsedTime(&a, c->ev[8], c->ev[0]) == hipSuccess &&
                hipEventElapsedTime(&b2, c->ev[9], c->ev[10]) == hipSuccess)
                c->timing.t[12] += 1e-3 * (a + b2);
        }
        const int* st = (const int*)(c->red_host + kRedStatus);
        if (st[0]) { capac_set_error("RATTLE_1 not converged!"); return CAPAC_ERR_RATTLE1; }
        if (st[1]) { capac_set_error("RATTLE_2 not converged!"); return CAPAC_ERR_RATTLE2; }
        ++done;
        if (save && stop_at_save) break;
    }
    return done;
}

extern "C" int capac_md_observables(capac_ctx* c, double* eksum, double* temp, double* ref_temp,
                                    double* vq, double* fmax, double* frms, double pot[CAPAC_NPOT],
                                    double vir[9]) {
    if (!c) return CAPAC_ERR_ARG;
    int rc = fetch_md(c);
    if (rc) return rc;
    const double* m = c->red_host + kRedMd;
    if (eksum) *eksum = m[0];
    if (temp) *temp = m[1];
    if (ref_temp) *ref_temp = c->cur_ref_temp;
    if (vq) memcpy(vq, m + 2, sizeof(double) * CAPAC_NHC);
    if (fmax) *fmax = m[23];
    if (frms) *frms = m[24];
    if (pot) {
        memcpy(pot, c->pot, sizeof c->pot);
        pot[0] = 0.0;
        for (int k = 1; k < CAPAC_NPOT; ++k) pot[0] += pot[k];    // sum_potential (force.c:322-331)
    }
    if (vir)
        for (int k = 0; k < 9; ++k) vir[k] = c->vir[k] + c->red_host[64 + k];
    return 0;
}

// ------------------------------------------------------------------ checkpoint

namespace {
// What the MD loop carries from one step to the next, beyond what traj.gro
// keeps (file.c:215-229 writes positions, velocities, vQ[0] and vP only).
struct CkHeader {
    char magic[8];                  // "CAPACCK2"
    int32_t n_atoms, n_metal, n_mat, step;
    int32_t solve_iters, rows_dirty, precision, pred_order, pred_count;
    int32_t bj_valid, bj_age;       // block preconditioner: formed, assemblies since
    int32_t reserved;
    double box[3], rcut, skin_req, skin;
    double first_temp, ext_temp, cur_ref_temp;
    double pot[CAPAC_NPOT], vir[9];
    double md[32];                  // kinetic energy, temperature, the whole NHC, F_max, F_rms, ...
    double red_md[16];              // RATTLE_2 virial of the last step
};
// then r, v, f, r_build (3N each), x and the predictor's three stored
// solutions, newest first (n_mat each), and the metal positions the block
// preconditioner's inverses were formed at (3 Nm, zeros when none)
const char kCkMagic[8] = {'C', 'A', 'P', 'A', 'C', 'C', 'K', '2'};
}  // namespace

extern "C" size_t capac_checkpoint_bytes(const capac_ctx* c) {
    if (!c) return 0;
    return sizeof(CkHeader) + sizeof(double) * (12 * (size_t)c->bc.N + 4 * (size_t)c->n_mat + 3 * (size_t)c->bc.Nm);
}

extern "C" int capac_checkpoint_save(capac_ctx* c, int step, void* buf, size_t bytes) {
    if (!c || !buf) return CAPAC_ERR_ARG;
    if (bytes < capac_checkpoint_bytes(c)) {
        capac_set_error("checkpoint buffer holds %zu bytes, %zu needed", bytes, capac_checkpoint_bytes(c));
        return CAPAC_ERR_ARG;
    }
    const size_t n3 = 3 * (size_t)c->bc.N;
    CkHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kCkMagic, 8);
    h.n_atoms = c->bc.N; h.n_metal = c->bc.Nm; h.n_mat = c->n_mat; h.step = step;
    h.solve_iters = c->solve_iters; h.rows_dirty = c->nb_dirty; h.precision = c->precision;
    h.pred_order = c->pred_order; h.pred_count = c->pred_count;
    h.bj_valid = c->bj_r && c->bj_valid; h.bj_age = c->bj_age;
    h.box[0] = c->bc.box[0]; h.box[1] = c->bc.box[1]; h.box[2] = c->bc.box[2];
    h.rcut = c->bc.rc; h.skin_req = c->skin_req; h.skin = c->skin;
    h.first_temp = c->first_temp; h.ext_temp = c->ext_temp; h.cur_ref_temp = c->cur_ref_temp;
    memcpy(h.pot, c->pot, sizeof h.pot);
    memcpy(h.vir, c->vir, sizeof h.vir);
    double* a = (double*)((char*)buf + sizeof(CkHeader));
    CHECK_HIP(hipMemcpyAsync(h.md, c->md, sizeof h.md, hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipMemcpyAsync(h.red_md, c->red + 64, sizeof h.red_md, hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipMemcpyAsync(a, c->r, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipMemcpyAsync(a + n3, c->v, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipMemcpyAsync(a + 2 * n3, c->f, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipMemcpyAsync(a + 3 * n3, c->r_build, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    for (int k = 0; k < 4 && c->n_mat; ++k)
        CHECK_HIP(hipMemcpyAsync(a + 4 * n3 + (size_t)k * c->n_mat, k ? c->x_hist[k - 1] : c->x,
                                 c->n_mat * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    double* bj_r = a + 4 * n3 + 4 * (size_t)c->n_mat;
    if (h.bj_valid)
        CHECK_HIP(hipMemcpyAsync(bj_r, c->bj_r, 3 * (size_t)c->bc.Nm * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    else
        memset(bj_r, 0, 3 * (size_t)c->bc.Nm * sizeof(double));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    memcpy(buf, &h, sizeof h);
    return 0;
}

extern "C" int capac_checkpoint_load(capac_ctx* c, const void* buf, size_t bytes, int* step) {
    if (!c || !buf) return CAPAC_ERR_ARG;
    CkHeader h;
    if (bytes < sizeof h) { capac_set_error("checkpoint is truncated"); return CAPAC_ERR_ARG; }
    memcpy(&h, buf, sizeof h);
    if (memcmp(h.magic, kCkMagic, 8) != 0) { capac_set_error("not a capac checkpoint"); return CAPAC_ERR_ARG; }
    if (h.n_atoms != c->bc.N || h.n_metal != c->bc.Nm || h.n_mat != c->n_mat || h.rcut != c->bc.rc ||
        h.box[0] != c->bc.box[0] || h.box[1] != c->bc.box[1] || h.box[2] != c->bc.box[2]) {
        capac_set_error("checkpoint of another system (%d atoms, %d metal, box %g %g %g)", h.n_atoms,
                        h.n_metal, h.box[0], h.box[1], h.box[2]);
        return CAPAC_ERR_ARG;
    }
    if (bytes < capac_checkpoint_bytes(c)) { capac_set_error("checkpoint is truncated"); return CAPAC_ERR_ARG; }
    int rc;
    if (h.skin_req != c->skin_req) {
        c->skin_req = h.skin_req;
        if ((rc = nb_regrid(c))) return rc;
    }
    const size_t n3 = 3 * (size_t)c->bc.N;
    const double* a = (const double*)((const char*)buf + sizeof(CkHeader));
    c->nb_dirty = 1;
    if (!h.rows_dirty) {
        // rebuild the neighbour rows at the positions they were built at, so
        // the rows (and the order of every pair sum) are those of the saved run
        CHECK_HIP(hipMemcpyAsync(c->r, a + 3 * n3, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        for (int attempt = 0;; ++attempt) {
            if ((rc = nb_build(c))) return rc;
            if ((rc = fetch_red(c))) return rc;
            if (!nb_overflowed(c)) break;
            if (attempt == 7) { capac_set_error("neighbour rows keep overflowing"); return CAPAC_ERR_HIP; }
            if ((rc = grow_nbr(c))) return rc;
        }
        c->nb_dirty = 0;
    }
    if (c->bj_inv) {
        // re-form the block preconditioner at the metal positions the saved
        // run formed it at, and take over its schedule
        c->bj_valid = 0;
        c->bj_age = 0;
        if (h.bj_valid) {
            const size_t Nm = (size_t)c->bc.Nm, N = (size_t)c->bc.N;
            const double* bj_r = a + 4 * n3 + 4 * (size_t)c->n_mat;
            for (int d = 0; d < 3; ++d)
                CHECK_HIP(hipMemcpyAsync(c->r + d * N, bj_r + d * Nm, Nm * sizeof(double), hipMemcpyHostToDevice,
                                         c->stream));
            if ((rc = cpim_assemble(c))) return rc;
            c->bj_age = h.bj_age;
        }
    }
    CHECK_HIP(hipMemcpyAsync(c->r, a, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CHECK_HIP(hipMemcpyAsync(c->v, a + n3, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CHECK_HIP(hipMemcpyAsync(c->f, a + 2 * n3, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    for (int k = 0; k < 4 && c->n_mat; ++k)
        CHECK_HIP(hipMemcpyAsync(k ? c->x_hist[k - 1] : c->x, a + 4 * n3 + (size_t)k * c->n_mat,
                                 c->n_mat * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CHECK_HIP(hipMemcpyAsync(c->md, h.md, sizeof h.md, hipMemcpyHostToDevice, c->stream));
    CHECK_HIP(hipMemcpyAsync(c->red + 64, h.red_md, sizeof h.red_md, hipMemcpyHostToDevice, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    c->solve_iters = h.solve_iters;
    c->batch_main = h.solve_iters;
    c->pred_order = h.pred_order;
    c->pred_count = h.pred_count; c->pred_pending = 0;
    c->first_temp = h.first_temp; c->ext_temp = h.ext_temp; c->cur_ref_temp = h.cur_ref_temp;
    memcpy(c->pot, h.pot, sizeof h.pot);
    memcpy(c->vir, h.vir, sizeof h.vir);
    if (step) *step = h.step;
    return fetch_md(c);
}

extern "C" int capac_set_thermostat(capac_ctx* c, const double* vq, double vp) {
    (void)vp;
    if (!c || !vq) return CAPAC_ERR_ARG;
    CHECK_HIP(hipMemcpyAsync(c->md + 2, vq, sizeof(double) * CAPAC_NHC, hipMemcpyHostToDevice, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int capac_time_kernel(capac_ctx* c, int kernel, int reps, double* ms) {
    if (!c || reps <= 0 || !ms) return CAPAC_ERR_ARG;
    if ((kernel == CAPAC_KERNEL_SPMV || kernel == CAPAC_KERNEL_ASSEMBLE || kernel == CAPAC_KERNEL_CPIM_MM ||
         kernel == CAPAC_KERNEL_APPLY64 || kernel == CAPAC_KERNEL_SOLVE) &&
        !c->n_mat) {
        capac_set_error("no CPIM unknowns");
        return CAPAC_ERR_ARG;
    }
    // the SpMV and the FP64 apply write into a solver work vector, the others
    // into their usual outputs; positions, b and the warm start are kept
    double* scratch = c->n_mat ? solver_scratch(c) : NULL;
    CHECK_HIP(hipStreamSynchronize(c->stream));
    CHECK_HIP(hipEventRecord(c->ev[14], c->stream));
    for (int k = 0; k < reps; ++k) {
        int rc = 0;
        switch (kernel) {
            case CAPAC_KERNEL_SPMV: rc = solver_spmv_only(c, c->x, scratch); break;
            case CAPAC_KERNEL_ASSEMBLE: rc = cpim_assemble(c); break;
            case CAPAC_KERNEL_PAIR: rc = pair_nonbonded(c, NULL); break;
            case CAPAC_KERNEL_NEIGHBOUR: rc = nb_build(c); break;
            case CAPAC_KERNEL_CPIM_MM: rc = cpim_mm_only(c); break;
            case CAPAC_KERNEL_APPLY64: rc = cpim_apply64(c, c->x, NULL, scratch, NULL, NULL); break;
            case CAPAC_KERNEL_SOLVE:
                // the last force call's system from the reference's warm
                // start: after MD steps the previous step's solution (second
                // newest stored), else the call's own start (x_backup); the
                // result lands in x as it did then
                rc = copy_d(c, c->pred_count >= 2 ? c->x_hist[1] : c->x_backup, c->x, c->n_mat);
                if (!rc) {
                    rc = solver_run(c, c->b, c->x);
                    if (rc == CAPAC_ERR_SOLVER) rc = 0;
                }
                break;
            default: capac_set_error("unknown kernel %d", kernel); return CAPAC_ERR_ARG;
        }
        if (rc) return rc;
    }
    CHECK_HIP(hipEventRecord(c->ev[15], c->stream));
    CHECK_HIP(hipEventSynchronize(c->ev[15]));
    float t = 0.f;
    CHECK_HIP(hipEventElapsedTime(&t, c->ev[14], c->ev[15]));
    *ms = (double)t / reps;
    return 0;
}

static int nb_build_count(capac_ctx* c, long long* n) {
    unsigned long long v = 0;
    CHECK_HIP(hipMemcpyAsync(&v, c->nb_ctl + 2, sizeof v, hipMemcpyDeviceToHost, c->stream));
    CHECK_HIP(hipStreamSynchronize(c->stream));
    *n = (long long)v;
    return 0;
}

extern "C" int capac_get_timing(capac_ctx* c, capac_timing* t) {
    if (!c || !t) return CAPAC_ERR_ARG;
    long long n = 0;
    int rc = nb_build_count(c, &n);
    if (rc) return rc;
    *t = c->timing;
    t->nb_builds = n - c->nb_builds_base;
    return 0;
}

extern "C" int capac_reset_timing(capac_ctx* c) {
    if (!c) return CAPAC_ERR_ARG;
    memset(&c->timing, 0, sizeof c->timing);
    return nb_build_count(c, &c->nb_builds_base);
}

extern "C" int capac_enable_timing(capac_ctx* c, int on) {
    if (!c) return CAPAC_ERR_ARG;
    c->timing_on = on;
    return 0;
}
