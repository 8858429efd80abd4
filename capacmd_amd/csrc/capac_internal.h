// Internal declarations of libcapac_gpu.so (HIP, gfx950).
//
// Device layout of one box (all FP64 unless noted):
//   r, v, f, old_r        3N SoA blocks x|y|z, the reference's rx/ry/rz layout
//   charge, inv_mass, mass, type(int), mol_id(int), local_idx(int), excl(uint)
//   cell list             cell_start[n_cell+1], cell_atoms[N] sorted by (cell, index)
//   neighbour rows        nbr[i * nb_stride + k]: non-metal neighbours from the
//                         front (k < n_front[i]), metal neighbours from the back
//                         (k >= nb_stride - n_back[i]); int32
//   relay (FP32-refined)  4 FP32 coefficients per metal pair: g_a = c2 - 3 c1,
//                         g_b = c1 r^2, c3, cq (relay.h); the separation is formed
//                         from the positions where a block is applied, and the
//                         diagonal blocks analytically.  Two layouts:
//                         rows   plane k of ordered pair (i,j) at
//                                relay32[(i*4 + k)*relay_ld + j];
//                         tiles  each unordered pair once: 64 x 64 tiles (I <= J)
//                                of metal blocks, tile t's pair (I*64 + l,
//                                J*64 + ((l + k) & 63)) as float4 at
//                                relay32[4 (t*4096 + k*64 + l)] (relay_tile.h).
//                         The FP64 path stores no relay: its product forms each
//                         block from the positions.
//   CPIM vector           x = [p (3Nm) | q (Nm) | lambda (n_np)], atomic units
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/capac_gpu.h"

#define CAPAC_WAVE 64

// reference literals (typedef.h:44,67,73,77; cpff.c:72,75,758,759)
#define K_INV_SQRT_PI 0.564189583547756
#define K_FQQ 138.935457696552
#define K_NM2BOHR 18.8972612456506
#define K_BOLTZ 0.00831446214546895
#define K_CONV_E 2.01552955019358e-05
#define K_CONV_V 3.80879884583368e-04
#define K_FQQ_AU2 49614.7526045429
#define K_FQQ_AU1 2625.49964037578

#define CHECK_HIP(x)                                                                    \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            capac_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return CAPAC_ERR_HIP;                                                       \
        }                                                                               \
    } while (0)

void capac_set_error(const char* fmt, ...);

// scalar constants shared by every kernel of a box
struct BoxConst {
    int N, Nm, n_np, n_types, n_mols;
    double box[3], half[3];
    double rc, rc2, alpha, a2_sqrtPI, wolfConst, erfc_arCut, inv_rc12, inv_rc6;
    double rl2;                 // neighbour-row radius squared: (rc + skin)^2
    int use_coulomb, use_vdw;
    int qsc_n, qsc_m;
    double qsc_eps, qsc_c, qsc_a;
    double inv_polar, inv_capac, inv_R_pp, inv_R_pq, inv_R_qq;
    double efield[3];
};

// coefficient planes per ordered metal pair of the FP32 relay (relay.h);
// its rows are allocated up to a multiple of kRelay32RowPad, the padding rows
// zero (the SpMV reads up to that many row-groups at once)
constexpr int kRelay32Planes = 4;
constexpr int kRelay32RowPad = 8;

struct PairParam {          // one entry of the n_types x n_types table
    int funct;                  // 1 LJ + Coulomb, 2 Buckingham, 3 Morse, 4 erf vdW
    double c6, c12, width, inv_width;
    double a, b;                // Buckingham A, B or Morse D, a (R in c6)
    // shifted-force LJ constants of the pair type (force.c:748-757):
    // f_c = 12 C12 / rc^12 - 6 C6 / rc^6, f_c / rc, and C12 / rc^12 - C6 / rc^6
    double lj_fc, lj_fc_rc, lj_ec;
};

// BiCGSTAB scalars, resident on the device (bicg_stab.c:289-292)
struct SolveScalars {
    double rho_s[2];        // rho of iteration n in slot n & 1
    double alpha, omega, bTb, criteria, rTr;
    int iter, done, kind;
};

struct Solver;              // solver.hip

// The device buffer `red` and its pinned host mirror share one layout, so a
// single transfer brings back everything the host reads:
//   [0, 256)    reduction outputs (energies, virials, momentum sums)
//   [256, 258)  neighbour-row flags (4 ints: overflow, spare)
//   [260, 292)  MD scalars (md)
//   [292, 294)  RATTLE failure flags (4 ints, md_status)
//   [296, 312)  solver scalars (SolveScalars)
//   [312, 316)  solver's FP64 residual norm
constexpr int kRedFlags = 256, kRedMd = 260, kRedStatus = 292, kRedSc = 296, kRedRes = 312;
constexpr int kRedWords = 316;
static_assert(sizeof(SolveScalars) <= (kRedRes - kRedSc) * sizeof(double), "solver scalars overflow their slot");

struct capac_ctx {
    int device, precision;
    BoxConst bc;
    int n_mat, n_cstr, ndf, fix_metal, nvt, n_heating;
    double dt, tau_t, ref_temp0;
    hipStream_t stream;

    // per atom
    double *r, *v, *f, *old_r;
    double *charge, *mass, *inv_mass;
    int *type, *mol_id, *local_idx;
    unsigned* excl;
    PairParam* pair;
    int pair_metal_inert;       // metal-metal pairs add nothing (uncharged metal, funct 1 with c6 = c12 = 0)
    int pair_lj_types;          // bit t: type t has funct-1 LJ (c6 or c12 nonzero) with a type the water kernel meets
    int pair_water;             // the water pair kernel applies (pair.hip): 1 to the non-metal atoms' front
                                // segments, 2 to every row (all pair types funct 1); then xyzq holds
    double* xyzq;               // its packed (x, y, z, q) copy, 4 doubles per atom
    // molecules / constraints / nanoparticles
    int *mol_first, *mol_last, *mol_cstr, *cstr_i, *cstr_j;
    double* cstr_r;
    int *np_first, *np_last, *np_of_metal;
    double* np_charge;
    // bonded terms and virtual sites, absolute atom indices (capac_params)
    int n_bonds, n_pairs14, n_angles, n_dihedrals, n_vsites;
    int *mol_terms, *bond_atoms, *bond_funct, *pair14_atoms, *angle_atoms, *angle_funct;
    int *dihedral_atoms, *vsite_atoms;
    double *bond_par, *pair14_par, *angle_par, *dihedral_par, *vsite_par;

    // cell list
    int ncell[3], n_cell;
    int *cell_of, *cell_count, *cell_start, *cell_atoms, *cell_tmp, *cell_fill, *cell_rank;
    double* sorted_r;           // 3N, positions in cell order
    int *sorted_mol, *sorted_loc;
    void* stencil;              // device copy of the cell stencil
    // neighbour rows
    int nb_stride;
    int *n_front, *n_back, *nbr, *nb_overflow;
    int* tail_list;             // [0] count, then the non-metal atoms with metal neighbours
    // Verlet skin: rows hold pairs within rc + skin; the device rebuilds them
    // when an atom has moved more than skin/2 since the last build
    double skin, skin_req;      // in use / requested (nm)
    int nb_dirty;               // positions or box replaced: build unconditionally
    long long nb_builds_base;   // build counter at the last timing reset
    double* r_build;            // 3N positions at the last build
    unsigned long long* nb_ctl; // [0] max displacement^2 bits, [1] rebuild flag, [2] builds,
                                // [3], [4] workgroup completion counts (zero between launches)

    // reductions: per-block partials and a small result area
    double* partial;            // per-block partials, one segment per producer (PartialSeg)
    int n_partial;
    int seg_off[8], seg_cap[8];  // segment offsets and capacities, in blocks of 16 doubles
    int n_pending;               // queued reductions, launched together by flush_reductions
    int pend_seg[8], pend_nblk[8], pend_nval[8], pend_red[8];
    double* red;                // kRedWords, laid out as described at kRedWords
    double* red_host;           // pinned mirror, same layout

    // Sutton-Chen
    double* isd;
    double* f_sc;               // its forces on the metal atoms (3 x N_m) when the CPIM force adds them
    double* mm_part;            // metal-metal CPIM force partials, Nb x Nb blocks of 3 x 64 (cpim.hip)
    // CPIM
    double *b, *x, *x_backup, *diag;
    int diag_set;               // diag holds the values for diag_polar / diag_capac
    double diag_polar, diag_capac;
    // solutions of the last MD steps, newest first, for the warm-start
    // predictor (order 0: the previous solution, as the reference)
    double* x_hist[3];
    int pred_order, pred_count;
    int pred_pending;           // extrapolation order the next force call applies
    float* relay32;
    int relay_layout;           // CAPAC_LAYOUT_ROWS or CAPAC_LAYOUT_TILES (FP32 path)
    int relay_nb;               // tiles: metal blocks of 64 per edge
    double* relay_part;         // tiles: the product's partial sums, Nb x Nb x 4 x 64
    int* relay_cnt;             // tiles: partials delivered per row block in the running product (Nb)
    double* metal_u;            // FP32 path: metal separations from metal atom 0 in bohr, 4 per atom
                                // (x, y, z, 0), then a flag at [4 Nm]: 1 when the metal is compact
    int relay_ld;
    size_t relay_bytes;
    // FP32 path: block-Jacobi preconditioner of GMRES (precond.hip), the
    // inverses of the relay's diagonal blocks of 64 metal atoms, re-formed
    // every bj_every assemblies (0: the Jacobi K instead)
    float* bj_inv;
    int bj_every, bj_age, bj_valid;
    double* bj_r;               // metal positions the inverses were formed at (x, y, z planes of Nm):
                                // a checkpoint carries them, so a restart re-forms the same inverses
    // its coarse space (cg_n = 4 per block of 64 metal atoms, 0 when off):
    // basis entries per metal atom (float4), active flags, Z^T q of the
    // running application, Ac^-1 (cg_n x cg_n)
    float* cg_w;
    int* cg_act;
    double *cg_s, *cg_A;
    int cg_n;
    Solver* solver;
    int solve_iters, solve_kind, solve_refine_rounds;
    int batch_main, batch_main2, batch_corr;  // iterations of the last two main solves and the last
                                              // correction solve (first batch sizes)
    double solve_crit;          // stop when rTr < solve_crit bTb (the reference: 1e-8)

    // fitting parameters (alpha, c, R_q, R_p, eps_SC, c_SC) and work space
    double theta[CAPAC_NTHETA];
    void* fit;
    void* em;                   // minimiser work space (em.hip)
    int fit_adjoint_iters;
    double solve_rTr, solve_bTb;

    // MD scalars (host copy of the small device state)
    double* md;                 // device: [eKsum, temp, vQ[10], aQ[10], scale, fmax2, frms]
    double first_temp, ext_temp, cur_ref_temp;
    int md_status_host;
    int* md_status;             // device RATTLE failure flags

    // results of the last force call
    double pot[CAPAC_NPOT], vir[9];
    // 0: the force call may leave the pair energies and virial unformed (an
    // MD step whose energies nobody reads); 1 otherwise (the default)
    int need_energy;
    // the last step's chain launch also ran the next step's first-half chain
    // and left its centre-of-mass pass to the next step's first pass
    int com_pending;
    // an MD step's force call copies its solution to hist_dst (the buffer
    // solver_push_history turns into the newest stored solution) in the CPIM
    // force launch, and records that buffer in hist_copied
    double* hist_dst;
    double* hist_copied;
    double pot_dev_ofs;         // unused
    // timing
    int timing_on;
    hipEvent_t ev[16];
    // second stream (default priority): pair and Sutton-Chen forces run on it
    // while the main stream assembles and solves the CPIM system
    hipStream_t side;
    hipEvent_t ev_rows, ev_side;
    capac_timing timing;
};

// ------------------------------------------------------------------ device helpers

// pbc_12 (force.c:111-121)
__device__ __forceinline__ void pbc12(const BoxConst& b, double& x, double& y, double& z) {
    // x < -half: x + L; x > half: x - L, formed as x - copysign(L, x) (one
    // magnitude test and one subtraction per component; x - (-L) rounds as
    // x + L, so the results are the same bits)
    x = fabs(x) > b.half[0] ? x - copysign(b.box[0], x) : x;
    y = fabs(y) > b.half[1] ? y - copysign(b.box[1], y) : y;
    z = fabs(z) > b.half[2] ? z - copysign(b.box[2], z) : z;
}

// 1/sqrt(s) for a positive finite s: the hardware estimate and one
// third-order Newton step, y + y e (1/2 + 3/8 e), e = 1 - s y^2 (the
// library rsqrt's own expansion without its zero and infinity cases)
__device__ __forceinline__ double rsqrt_pos(double s) {
    const double y = __builtin_amdgcn_rsq(s);
    const double e = fma(-s * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// dist_2 (force.c:41-44) evaluated exactly as the reference: (x*x + y*y) + z*z,
// each product and sum rounded once, no fused multiply-add
__device__ __forceinline__ double dist2_exact(double x, double y, double z) {
#pragma clang fp contract(off)
    double xx = x * x;
    double yy = y * y;
    double zz = z * z;
    return (xx + yy) + zz;
}

// wavefront sum of a double, fixed butterfly order (deterministic)
// Kernels on the main stream that run while the pair kernel occupies the
// same compute units (the CPIM right-hand side, the assembly and the solve,
// a chain of short latency-bound launches) raise their waves' issue
// priority, so that a SIMD shared with pair-kernel waves serves theirs first.
// This is the waves' priority inside a kernel; the streams keep the default.
__device__ __forceinline__ void main_stream_priority() { __builtin_amdgcn_s_setprio(3); }

// Wavefront sum through DPP lane moves (VALU) instead of LDS-crossbar
// shuffles: quad sums, 8- and 16-lane sums by mirrors, then the row
// broadcasts gather the four rows into lane 63, which every lane reads.
// The order of the additions is fixed, so results are reproducible.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_move<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_move<0x141, 0xF>(v);   // row_half_mirror: 8-lane sums
    v += dpp_move<0x140, 0xF>(v);   // row_mirror: 16-lane sums
    v += dpp_move<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// The fixed-order sums of nval (<= 16) values over nblk partial rows of 16
// doubles, by one whole workgroup (blockDim.x a multiple of 64, at most
// 1024): thread t adds rows t, t + blockDim.x, ... in turn, the wavefronts
// sum their threads, and the wavefronts' sums are added in order.  Threads
// t < nval write the sum to out[t] and to s_out[t] (LDS).
__device__ __forceinline__ void reduce_rows_block(const double* __restrict__ partial, int nblk, int nval,
                                                  double* __restrict__ out, double (*s_w)[16], double* s_out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] += partial[b * 16 + v];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        double s = wave_sum(acc[v]);
        if (lane == 0) s_w[wave][v] = s;
    }
    __syncthreads();
    if (threadIdx.x < nval) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_w[w][threadIdx.x];
        out[threadIdx.x] = t;
        if (s_out) s_out[threadIdx.x] = t;
    }
}

// ------------------------------------------------------------------ launchers

// neighbour.hip
int nb_grid_setup(capac_ctx* c);
int nb_build(capac_ctx* c);                 // unconditional build at rc + skin
int nb_build_if_needed(capac_ctx* c);       // device-side decision from displacements
int nb_set_radius(capac_ctx* c, double skin);
// pair.hip
int pair_nonbonded(capac_ctx* c, double* f_out_accumulate, int max_blocks = 0, int field = 0);
// sc.hip
// f_sc NULL: adds the forces to c->f; else writes them to f_sc (3 x N_m)
int sc_forces(capac_ctx* c, double* f_sc = NULL);
// cpim.hip
int cpim_rhs(capac_ctx* c);
int cpim_assemble(capac_ctx* c);
int cpim_rhs_assemble(capac_ctx* c);       // both, in one launch where the layout allows
// whether a force call's Sutton-Chen passes run inside its RHS/relay and CPIM
// force launches (the FP32 path's TILES layout) instead of launches of their own
bool sc_in_force_launches(const capac_ctx* c);
int cpim_forces(capac_ctx* c);
// f_sc (3 x N_m, or NULL): forces the metal-metal launch adds with its own
// (on the FP32 path's TILES layout the same launch writes them first)
int cpim_forces_x(capac_ctx* c, const double* x, double* f_sc = NULL);
int cpim_dense_expand(capac_ctx* c, double* dense_dev);
// out = b - A x (b non-NULL) or A x, with the FP64 relay recomputed on the fly
// (no stored matrix); part[g] = (out, out) over the rows of row-group g,
// *n_part of them (row-groups: metal atoms, or blocks of 64 on the TILES
// layout, then the Lagrange rows).  With a non-NULL guard the launches do
// nothing unless that solve has finished with kind converged or loosened.
int cpim_apply64(capac_ctx* c, const double* x, const double* b, double* out, double* part,
                 const SolveScalars* guard, int* n_part = NULL);
// precond.hip
size_t bj_bytes(int Nm);                    // the inverses' size
int bj_alloc_dev(capac_ctx* c);             // (re)allocates for bj_every (none when 0)
void bj_free(capac_ctx* c);
int bj_build(capac_ctx* c);                 // inverts the diagonal blocks of the FP32 relay
int bj_copy_block(capac_ctx* c, int b, float* out);   // block b's inverse, row-major, host
int bj_copy_coarse(capac_ctx* c, int* n, double* ainv, float* basis);   // coarse space, host
// The preconditioner M^-1 = B^-1 + Z Ac^-1 Z^T in two launches: bj_apply
// sets out = B^-1 q (or adds it: add) and keeps Z^T q in cg_s; cg_apply adds
// Z Ac^-1 cg_s to out (nothing without a coarse space).  guard 1: no-op
// once sc reports done, 2: no-op until then.
void bj_apply(capac_ctx* c, const double* q, double* out, bool add, const SolveScalars* sc, int guard);
void cg_apply(capac_ctx* c, double* out, const SolveScalars* sc, int guard);
// solver.hip
int solver_create(capac_ctx* c);
void solver_destroy(capac_ctx* c);
int solver_run(capac_ctx* c, const double* b_dev, double* x_dev);
int solver_noalloc_init(int device);        // device address of the allocation-free work space
int solver_run_noalloc(capac_ctx* c, const double* b_dev, double* x_dev);
int solver_spmv_only(capac_ctx* c, const double* x, double* y);
double* solver_scratch(capac_ctx* c);
void solver_predict(capac_ctx* c);          // x from the history (md steps), at the next force call
int force_prologue(capac_ctx* c, int n_red); // red cleared, pending extrapolation, x_backup
void solver_push_history(capac_ctx* c);     // the solution just found joins it
int cpim_mm_only(capac_ctx* c);
// bonded.hip
bool has_bonded(const capac_ctx* c);        // bonds, 1-4 pairs, angles or dihedrals
int bonded_blocks(int n_mols);              // workgroups of the bonded kernel
int bonded_forces(capac_ctx* c);            // adds them to c->f; red[112..126)
int vsite_build(capac_ctx* c);              // site positions from their atoms
int vsite_spread(capac_ctx* c, double* f);  // site forces onto their atoms
// em.hip
void em_destroy(capac_ctx* c);
int em_begin(capac_ctx* c);
int em_move(capac_ctx* c, double length);
int em_direction(capac_ctx* c, int cg);
// fit.hip
void fit_destroy(capac_ctx* c);
void fit_apply_theta(capac_ctx* c);
int fit_frame_loss_grad(capac_ctx* c, const double* f_ref, double e_ref, double w_e, double* loss,
                        double* grad);
// md.hip
int md_kick(capac_ctx* c);
int md_rattle1(capac_ctx* c);
int md_rattle2(capac_ctx* c, bool with_kick = false);   // with_kick: md_kick first (fused with constraints)
int md_first_half(capac_ctx* c, double kT);    // chain, scale, half kick, drift
// kinetic energy, chain, scale, COM removal; kT_next > 0: an unobserved step
// whose next step follows in the same call (see md.hip)
int md_second_half(capac_ctx* c, double kT, double kT_next = 0.0);
int md_remove_com_kinetic(capac_ctx* c);   // after the first force call
int md_wrap(capac_ctx* c);
int md_fmax(capac_ctx* c);
// reductions (api.hip): sum per-block partials [nblk x 16] into red[0..nval)
// Per-block partial sums [nblk x 16] are written into the producer's own
// segment of c->partial; reduce_partials queues their fixed-order reduction
// into red[red_offset..+nval), and flush_reductions launches every queued one
// in a single kernel (before anything reads red).
enum PartialSeg { SEG_PAIR, SEG_SC, SEG_MW, SEG_MM, SEG_RATTLE, SEG_MOM, SEG_BONDED, SEG_COUNT };
double* partial_seg(capac_ctx* c, int seg);
int reduce_partials(capac_ctx* c, int seg, int nblk, int nval, int red_offset);
// the same for sums that only energies and the virial need: nothing when the
// force call may leave them unformed (need_energy 0)
int reduce_energies(capac_ctx* c, int seg, int nblk, int nval, int red_offset);
int flush_reductions(capac_ctx* c);
int zero_outputs(capac_ctx* c);
// graph-safe clears (fill kernels, not hipMemsetAsync)
int fill_zero_i(capac_ctx* c, int* p, size_t n);
int fill_zero_d(capac_ctx* c, double* p, size_t n);
