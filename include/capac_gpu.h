/* capac_gpu.h -- C ABI of the MI355X force + CPIM path (libcapac_gpu.so).
 *
 * One context owns one simulation box resident in HBM.  The entry points
 * replace the reference's per-step operator
 *
 *   void mpi_force(Task*, Topol*, Atom_Info*, Mol_Info*, RunSet*, Metal*,
 *                  System*, Bicgstab*, int my_id, int num_procs,
 *                  double** time_used)            (reference force.h:100,
 *                                                   force.c:2097-2320)
 *
 * and its sub-operators (cpff.h:30,38,45, bicg_stab.h:58, force.c:385-542,
 * 1402-1540), plus the MD step body that the reference runs on rank 0
 * (mpi-main.c:874-1013).  Plain pointers and sizes only.
 *
 * Conventions, as the reference:
 *   - MD units: nm, ps, kJ/mol, e, amu.  CPIM vectors in atomic units.
 *   - Per-atom 3-vectors are SoA blocks x[N] | y[N] | z[N] (the reference's
 *     rx/ry/rz and fx/fy/fz layout, mpi-main.c:476-488).
 *   - pot[15] uses the reference's index map (mpi-main.c:431-446):
 *     1 Sutton-Chen, 7 Coulomb, 8 vdW, 9 external field, 10-14 CPIM terms,
 *     0 = sum of 1..14 (force.c:322-331).
 *   - vir[9] is the reference's virial[3][3], row-major.
 *   - The CPIM vector x = vec_pq is [p (3 N_m, interleaved xyz) | q (N_m) |
 *     lambda (n_NP)] in atomic units (cpff.c:167-193).
 *   - Metal atoms must be atoms [0, n_metal) (cpff.c:762-764).
 * Errors are returned, never exit()ed: 0 = success, < 0 = failure; the
 * message is in capac_last_error().  One host thread per context; calls are
 * synchronous at the API boundary.
 */
#ifndef CAPAC_GPU_H
#define CAPAC_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAPAC_NPOT 15
#define CAPAC_NHC 10              /* Nose-Hoover chain length (mpi-main.c:514) */

/* return / status codes */
#define CAPAC_OK               0
#define CAPAC_ERR_ARG        (-1)
#define CAPAC_ERR_HIP        (-2)
#define CAPAC_ERR_SOLVER     (-3)  /* BiCGSTAB not converged (bicg_stab.c:563-569) */
#define CAPAC_ERR_RATTLE1    (-4)  /* RATTLE_1 not converged (force.c:1990-1994) */
#define CAPAC_ERR_RATTLE2    (-5)  /* RATTLE_2 not converged (force.c:2083-2087) */
#define CAPAC_ERR_UNSUPPORTED (-6)

/* precision of the CPIM relay matrix and inner solve */
#define CAPAC_PREC_FP64          0  /* FP64 matrix, FP64 BiCGSTAB (the reference's arithmetic) */
/* FP32 relay (T-tensors evaluated in FP32, 4 coefficients per metal pair,
 * the separation formed from the positions where a block is applied),
 * GMRES on it (restarted every 40 columns, right-preconditioned with the
 * reference's Jacobi preconditioner) with FP64 vectors and dots, and FP64
 * iterative refinement: the reference's stop test is applied to the true
 * residual b - A x with the FP64 relay recomputed on the fly, and
 * corrections are solved for until it holds.  Iteration counts of this path
 * are GMRES columns (one product each; a BiCGSTAB iteration takes two).  No
 * counterpart in the reference (its arithmetic is FP64 only). */
#define CAPAC_PREC_FP32_REFINED  1

/* Verlet skin of the neighbour rows (nm), see capac_set_skin */
#define CAPAC_DEFAULT_SKIN 0.1

/* BiCGSTAB stop test rTr < crit bTb; the reference's crit (bicg_stab.c:303-306) */
#define CAPAC_DEFAULT_SOLVE_CRIT 1.0e-8

/* solver exit kinds (bicg_stab.c:300,368,539-571) */
#define CAPAC_SOLVE_CONVERGED   0   /* rTr < 1e-8 bTb */
#define CAPAC_SOLVE_LOOSENED    1   /* iter > 1000 and rTr < 1e-6 bTb */
#define CAPAC_SOLVE_FAILED      2   /* iter > 1000 and rTr >= 1e-6 bTb */
#define CAPAC_SOLVE_ZERO_RHS    3   /* bTb < 1e-10: x left unchanged */
#define CAPAC_SOLVE_ZERO_RHO    4   /* (r0,r) < 1e-10 at the first iteration */

/* Everything the path reads from param.txt/mdset.txt, flattened.  The
 * pointers must stay valid only for the duration of capac_create(). */
typedef struct capac_params {
    int n_atoms;
    int n_metal;                /* metal atoms are [0, n_metal)              */
    int n_nps;                  /* nanoparticles                             */
    int n_types;                /* vdW types                                 */
    int n_mols;                 /* molecules                                 */
    int n_cstr;                 /* constraints, all molecules                */

    /* per atom */
    const double*   charge;     /* e                                         */
    const double*   mass;       /* amu (0 for virtual sites)                 */
    const int*      atomtype;   /* row/column of the pair table              */
    const int*      mol_id;     /* molecule index in the system              */
    const int*      local_idx;  /* index inside its molecule (< 32)          */
    const unsigned* excl_mask;  /* bit k: excluded from the atom with local
                                   index k of the same molecule
                                   (force.c:1457-1461)                       */
    /* per molecule */
    const int*      mol_first;  /* first atom                                */
    const int*      mol_last;   /* last atom                                 */
    const int*      mol_cstr;   /* constraint range [mol_cstr[m], mol_cstr[m+1]) */
    /* constraints, absolute atom indices, in param.txt order per molecule */
    const int*      cstr_i;
    const int*      cstr_j;
    const double*   cstr_r;     /* nm                                        */
    /* nanoparticles: metal index ranges and total charges */
    const int*      np_first;
    const int*      np_last;
    const double*   np_charge;
    /* pair table, n_types x n_types row-major, symmetric (file.c:618-695);
       funct 1 = LJ (C6, C12) + Coulomb, 2 = Buckingham (A, B in nb_a, nb_b;
       C6), 3 = Morse (D, a in nb_a, nb_b; R in nb_c6), 4 = erf-damped vdW
       (C6, C12, width in nb_a); funct 2-4 carry no Coulomb term
       (force.c:1476-1525) */
    const int*      nb_funct;
    const double*   nb_c6;
    const double*   nb_c12;
    const double*   nb_a;
    const double*   nb_b;

    double box[3];              /* nm, rectangular                           */
    double rcut;                /* nm                                        */
    double dsf_alpha;           /* nm^-1                                     */
    int    use_coulomb;         /* 1 = damped shifted force (wolf_sum), 0 = plain cut-off */
    int    use_vdw;             /* 1 = shifted-force LJ, 0 = plain cut-off   */
    int    qsc_n, qsc_m;
    double qsc_eps, qsc_c, qsc_a;   /* kJ/mol, -, nm                         */
    int    use_cpff;
    double cpff_polar;          /* bohr^3                                    */
    double cpff_capac;          /* bohr                                      */
    double efield[3];           /* kJ/mol/(e nm)                             */

    /* integrator */
    double dt;                  /* ps                                        */
    int    fix_metal;
    int    nvt;                 /* 1: Nose-Hoover chain (ensemble nvt)       */
    double ref_temp;            /* K                                         */
    double tau_t;               /* ps                                        */
    int    n_heating;
    int    ndf;                 /* 3(N - n_vsites) - 3 - n_cstr (mpi-main.c:548) */

    /* Bonded terms and virtual sites (mpi_bonded force.c:1694-1904,
     * build_vsites / mpi_vsites force.c:1582-1692), expanded to every
     * molecule with absolute atom indices; within a molecule each list keeps
     * param.txt order.  mol_terms holds five ranges of n_mols + 1 entries,
     * [CAPAC_TERM_BOND .. CAPAC_TERM_VSITE], term t of molecule m running
     * over [mol_terms[t (n_mols + 1) + m], mol_terms[t (n_mols + 1) + m + 1]).
     * All counts 0 and pointers NULL for rigid molecules and metal only. */
    int n_bonds, n_pairs14, n_angles, n_dihedrals, n_vsites;
    const int*    mol_terms;
    const int*    bond_atoms;     /* i, j                                         */
    const int*    bond_funct;     /* 1 harmonic, 3 Morse (funct 5 bonds only exclude) */
    const double* bond_par;       /* b0, kb (funct 1) or b0, D, beta (funct 3)    */
    const int*    pair14_atoms;   /* i, j                                         */
    const double* pair14_par;     /* c6, c12, q_i, q_j, scale_lj, scale_qq: funct 1
                                     takes the pair table and the SCALE_LJ_QQ
                                     line, funct 2 its own line (force.c:1752-1790) */
    const int*    angle_atoms;    /* i, j, k                                      */
    const int*    angle_funct;    /* 1 harmonic, 5 Urey-Bradley                   */
    const double* angle_par;      /* a0 (degrees), ka, b0_ub, kb_ub               */
    const int*    dihedral_atoms; /* i, j, k, l                                   */
    const double* dihedral_par;   /* Ryckaert-Bellemans c0..c5; periodic dihedrals
                                     (funct 1, 4, 9) arrive converted as
                                     compute_dihedral_9 converts them (force.c:1236-1347) */
    const int*    vsite_atoms;    /* i, j, k, s (the site)                        */
    const double* vsite_par;      /* a, b, c of the TIP5P-type construction       */
} capac_params;

#define CAPAC_TERM_BOND     0
#define CAPAC_TERM_PAIR14   1
#define CAPAC_TERM_ANGLE    2
#define CAPAC_TERM_DIHEDRAL 3
#define CAPAC_TERM_VSITE    4

/* per-phase wall times in seconds, the reference's timer categories
 * (mpi-main.c:119-128): [1] SC density [2] SC force [4] nonbonded
 * [5] CPIM matrix [6] CPIM vector [7] CPIM solve [8] CPIM force
 * [10] neighbour build [12] integrator [0] total force */
typedef struct capac_timing {
    double t[CAPAC_NPOT];
    long long n_force_calls;
    long long solve_iters;      /* BiCGSTAB iterations, summed               */
    long long nb_builds;        /* neighbour-row builds (see capac_set_skin) */
    long long solve_syncs;      /* host synchronisations inside the solves   */
} capac_timing;

typedef struct capac_ctx capac_ctx;

const char* capac_last_error(void);
const char* capac_version(void);

/* Device set-up: copies the topology and parameters to HBM, sizes the
 * neighbour list, relay matrix and solver work space.  Replaces the
 * reference's allocation block (mpi-main.c:448-609). */
int  capac_create(const capac_params* p, int precision, int device, capac_ctx** out);
void capac_destroy(capac_ctx* ctx);

/* Positions/velocities (3N SoA, nullable) and the CPIM vector (n_mat, nullable,
 * the solver's warm start, mpi-main.c:585-590). */
int capac_set_state(capac_ctx* ctx, const double* r, const double* v, const double* vec_pq);
int capac_get_state(capac_ctx* ctx, double* r, double* v, double* vec_pq);
int capac_get_forces(capac_ctx* ctx, double* f);
int capac_set_box(capac_ctx* ctx, const double box[3]);

/* Verlet skin of the neighbour rows, nm (no counterpart in the reference,
 * which scans all pairs at every call).  Rows hold the pairs within
 * rc + skin and are rebuilt, on the device, once an atom has moved more than
 * skin/2 since the last build; every force term still applies the exact
 * r^2 < rc^2 test, so results do not depend on the skin.  0 rebuilds at every
 * force call.  A box edge shorter than 2 (rc + skin) turns the skin off.
 * capac_skin returns the skin in use. */
int capac_set_skin(capac_ctx* ctx, double skin);
double capac_skin(const capac_ctx* ctx);

/* Stop test of the CPIM solve, rTr < crit bTb (and the 1000-iteration
 * loosening at 100 crit); default CAPAC_DEFAULT_SOLVE_CRIT, the reference's
 * fixed value.  No counterpart in the reference: it exists so that accuracy
 * studies can solve both precisions far below the reference's tolerance. */
int capac_set_solve_tolerance(capac_ctx* ctx, double crit);

/* Warm start of each MD step's CPIM solve (capac_md_run).  Order 0 starts
 * from the previous step's solution, as the reference does
 * (bicg_stab.c:303-306); order 1 extrapolates the last two solutions
 * linearly, order 2 the last three quadratically, which takes far fewer
 * iterations to the same stop test.  Defaults: 0 for CAPAC_PREC_FP64 (the
 * reference's iteration path), CAPAC_DEFAULT_PREDICTOR_FP32 for
 * CAPAC_PREC_FP32_REFINED.  capac_md_begin starts the stored solutions;
 * capac_set_state with positions or a solution, and capac_set_box, clear
 * them.  No counterpart in the reference. */
int capac_set_predictor(capac_ctx* ctx, int order);
int capac_predictor(const capac_ctx* ctx);

/* Storage layout of the FP32 relay (CAPAC_PREC_FP32_REFINED only).  ROWS
 * keeps the 4 coefficients of every ordered metal pair in row-major planes
 * (16 B per ordered pair); TILES keeps each unordered pair once in 64 x 64
 * tiles of metal blocks (8 B per ordered pair) and its product applies every
 * stored pair to both of its rows, with fixed-order partial sums, so it
 * moves half the bytes.  AUTO (the default at capac_create) takes TILES from
 * CAPAC_TILES_MIN_METAL metal atoms on, where the tiles fill the chip.  Both
 * hold the same FP32 coefficients; the products differ by FP64 summation
 * order only.  Changing the layout reallocates the relay; the next force
 * call assembles it.  capac_fp32_layout returns the layout in use (-1 on
 * the FP64 path).  No counterpart in the reference. */
#define CAPAC_LAYOUT_AUTO   0
#define CAPAC_LAYOUT_ROWS   1
#define CAPAC_LAYOUT_TILES  2
#define CAPAC_TILES_MIN_METAL 1024
int capac_set_fp32_layout(capac_ctx* ctx, int layout);
int capac_fp32_layout(const capac_ctx* ctx);

/* Preconditioner of the FP32-refined path's GMRES.  every > 0: block Jacobi,
 * the exact inverses (FP32) of the relay's diagonal blocks of 64 metal
 * atoms (256 unknowns each: their dipoles and charges), plus an additive
 * coarse correction over four charge vectors per block (1, x, y, z over its
 * atoms, orthonormalised; up to 64 blocks), all formed at the first
 * assembly and again every `every` assemblies (an MD step assembles once);
 * every = 0: the reference's Jacobi diagonal.  The preconditioner steers
 * the iterations only: the stop test, and so the accuracy of a solution, is
 * the same.  The inverses take 256 KiB per block of 64 metal atoms.
 * capac_block_precond returns the interval in use (-1 on the FP64 path).
 * No counterpart in the reference. */
#define CAPAC_BJ_EVERY_DEFAULT 32
int capac_set_block_precond(capac_ctx* ctx, int every);
int capac_block_precond(const capac_ctx* ctx);
/* Diagnostic: the inverse of diagonal block `block` (0 .. ceil(Nm/64) - 1)
 * as last formed, row-major 256 x 256 into out (host).  Unknown rho < 192 is
 * dipole component rho % 3 of metal atom 64 block + rho / 3, rho >= 192 the
 * charge of metal atom 64 block + rho - 192; unknowns past Nm are identity
 * rows.  No counterpart in the reference. */
int capac_block_precond_inverse(capac_ctx* ctx, int block, float* out);
/* Diagnostic: the coarse space as last formed.  *n = its size (4 per block;
 * 0 when it is off or not formed yet); ainv (n x n, row-major) receives
 * Ac^-1, basis (4 floats per atom slot, ceil(Nm/64) * 64 slots) the basis
 * entries over the metal charges, a dropped vector's entries zero.  Either
 * may be NULL.  No counterpart in the reference. */
int capac_block_precond_coarse(capac_ctx* ctx, int* n, double* ainv, float* basis);

/* ---- restart checkpoint (no counterpart in the reference) ----
 * The reference restarts from the last traj.gro frame, which keeps
 * positions, velocities, vQ[0] and vP only (file.c:152-160, 215-229): the
 * other nine Nose-Hoover links and the CPIM warm start vec_pq are lost, so
 * a restarted run leaves the trajectory it continues.  A checkpoint holds
 * everything the MD loop carries between steps (positions, velocities,
 * forces, vec_pq, the whole chain, the heating ramp, the neighbour-row
 * build positions, the predictor's stored solutions and order, the metal
 * positions the block preconditioner was formed at and its schedule) so that
 * capac_checkpoint_load followed by capac_md_run
 * continues the saved run step for step, bit for bit on the same device and
 * precision.  Save between capac_md_run calls; `step` is stored as given.
 * Loading checks that the checkpoint belongs to this system (atom counts,
 * box, cut-off) and takes over its Verlet skin.  Minimiser state is not
 * included. */
size_t capac_checkpoint_bytes(const capac_ctx* ctx);
int capac_checkpoint_save(capac_ctx* ctx, int step, void* buf, size_t bytes);
int capac_checkpoint_load(capac_ctx* ctx, const void* buf, size_t bytes, int* step);

/* ---- energy minimisation, run_type em / cg (mpi-main.c:706-865) ----
 * capac_em_begin after the first capac_force: direction s = F, F_old = F.
 * capac_em_step: r_old = r; r += s / F_max * length (metal stays with
 * fix_metal); RATTLE_1 on the positions; v = 0; one force call; returns the
 * potential energies with pot[0] = sum of pot[1..14], F_max and F_rms.
 * capac_em_direction (only when not converged): s = F + gamma s with
 * gamma = 0 (cg 0, steepest descent) or (F.F - F_old.F)/(F_old.F_old)
 * (cg 1), then F_old = F.  The step-length rule and the convergence test are
 * the caller's (the CLI follows mpi-main.c:781-819). */
int capac_em_begin(capac_ctx* ctx);
int capac_em_step(capac_ctx* ctx, double length, double pot[CAPAC_NPOT], double* fmax, double* frms);
int capac_em_direction(capac_ctx* ctx, int cg);

/* ---- fitting (no counterpart in the reference) ----
 * theta = (alpha [bohr^3], c [bohr], R_q [bohr], R_p [bohr], eps_SC [kJ/mol],
 * c_SC): CPIM polarizability and capacitance, the two Gaussian widths as
 * free parameters (the reference derives them from alpha and c,
 * mpi-main.c:318-323; capac_create starts from those derived values) and
 * the Sutton-Chen energy scale and embedding constant.  Setting theta
 * changes every later force call. */
#define CAPAC_NTHETA 6
#define CAPAC_DEFAULT_PREDICTOR_FP32 2
int capac_set_theta(capac_ctx* ctx, const double theta[CAPAC_NTHETA]);
int capac_get_theta(capac_ctx* ctx, double theta[CAPAC_NTHETA]);

/* One frame of the force-matching loss at the resident state (positions and
 * the warm-start vec_pq set by the caller):
 *   L = sum_{metal i} |F_i - f_ref_i|^2 + w_e (E - e_ref)^2,
 * F from one capac_force call, E = sum of pot[1..14], f_ref 3 N_m SoA
 * (x block, y block, z block).  grad (nullable) receives dL/dtheta from the
 * adjoint method: A lambda = dL/dx solved with the BiCGSTAB kernels (A is
 * symmetric), forward-mode derivative kernels that recompute the T-tensors
 * for dA/dtheta, dF/dtheta and dE/dtheta, and dL/dx from the derivative of
 * dE/dx along G = 2 (F - f_ref) (the CPIM forces are -dE/dr at fixed x). */
int capac_fit_frame(capac_ctx* ctx, const double* f_ref, double e_ref, double w_e, double* loss,
                    double grad[CAPAC_NTHETA]);

/* One force evaluation = mpi_force (force.c:2097-2320): zeroes, pair forces,
 * Sutton-Chen, CPIM right-hand side, relay assembly, BiCGSTAB (warm start
 * from the resident vec_pq), CPIM forces and the external field.  pot[0] is
 * left as the reference leaves it (the caller sums); f (3N) and the outputs
 * are nullable.  Returns CAPAC_ERR_SOLVER where the reference exits. */
int capac_force(capac_ctx* ctx, double pot[CAPAC_NPOT], double vir[9], double* f);

/* ---- sub-operators, each one kernel family, for parity tests ---- */

/* Neighbour list (cell list + the reference's exact cut-off test):
 * per atom i, every j != i not excluded with |pbc_12(r_j - r_i)|^2 < rl^2,
 * rl = rc + capac_skin(ctx), built now.  When the skin is 0, rl^2 is rc^2
 * widened by a relative 1e-12, so that no pair within rounding of rc is
 * missed (every force term applies the exact r^2 < rc^2 test itself).
 * count[N]; list[N * max_per_atom] (j indices, unordered); returns the
 * largest row length, or < 0. */
int capac_neighbours(capac_ctx* ctx, int* count, int* list, int max_per_atom);

/* mpi_nonb_pair (force.c:1402-1540): f, pot[7], pot[8], vir (each nullable;
 * values are written, not accumulated) */
int capac_nonbonded(capac_ctx* ctx, double* f, double pot[CAPAC_NPOT], double vir[9]);

/* mpi_bonded (force.c:1694-1904) at the resident positions: bonds, 1-4
 * pairs, angles and dihedrals.  f, pot[2], pot[3], pot[4], and the 1-4
 * pairs' pot[7], pot[8], vir (each nullable; written, not accumulated) */
int capac_bonded(capac_ctx* ctx, double* f, double pot[CAPAC_NPOT], double vir[9]);

/* build_vsites (force.c:1582-1629) on the resident positions, which it
 * updates; r (3N, nullable) receives them */
int capac_vsite_build(capac_ctx* ctx, double* r);

/* mpi_vsites (force.c:1636-1692): moves the forces f_in (3N) on the virtual
 * sites to their constructing atoms at the resident positions; f_out (3N) */
int capac_vsite_spread(capac_ctx* ctx, const double* f_in, double* f_out);

/* mpi_qsc_dens + mpi_qsc_force (force.c:385-542) */
int capac_sutton_chen(capac_ctx* ctx, double* f, double pot[CAPAC_NPOT], double vir[9],
                      double* inv_sqrt_dens);

/* mpi_cpff_vec_ext (cpff.c:43-198): b[n_mat] */
int capac_cpim_rhs(capac_ctx* ctx, double* b);

/* mpi_cpff_mat_relay_COO (cpff.c:205-519), assembled into the device layout;
 * optionally expanded to the dense (4 N_m) x (4 N_m) block without the
 * Lagrange rows (row-major, nullable) and the Jacobi diagonal diag[n_mat]. */
int capac_cpim_assemble(capac_ctx* ctx, double* dense, double* diag);

/* mpi_precon_bicg_stab_COO (bicg_stab.c:253-573) on the assembled matrix.
 * b (nullable: the resident right-hand side), x in/out (nullable: the
 * resident vec_pq).  kind = CAPAC_SOLVE_*. */
int capac_cpim_solve(capac_ctx* ctx, const double* b, double* x, int* iters,
                     double* rTr, double* bTb, int* kind);

/* capac_cpim_solve without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it).  The solver's work vectors,
 * partial sums and scalars live in the library's file-scope device array
 * (about 2.2 MB per device, holding systems of at most 16384 CPIM unknowns;
 * larger systems are refused with CAPAC_ERR_ARG) instead of the context's
 * allocations; the call itself allocates nothing.  Same arguments, kernels
 * and results as capac_cpim_solve.  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU race. */
int capac_cpim_solve_noalloc(capac_ctx* ctx, const double* b, double* x, int* iters,
                             double* rTr, double* bTb, int* kind);

/* mpi_cpff_force (cpff.c:526-950) for the given x (nullable: resident) */
int capac_cpim_force(capac_ctx* ctx, const double* x, double* f, double pot[CAPAC_NPOT],
                     double vir[9]);

/* statistics of the last solve */
int capac_solve_stats(capac_ctx* ctx, int* iters, double* rTr, double* bTb, int* kind);

/* ---- MD on the device (mpi-main.c:629-640, 874-1013) ---- */

/* After the first force call: remove the centre-of-mass motion and compute
 * the kinetic energy; records the starting temperature. */
int capac_md_begin(capac_ctx* ctx);

/* Runs steps first_step .. first_step + n_steps - 1 of the MD loop: NHC half
 * step, velocity Verlet with RATTLE, force, NHC, centre-of-mass removal, and
 * the molecule wrap every n_save steps (force.c:127-165).  Stops early after
 * a step that is a multiple of n_save when stop_at_save is set, so the host
 * can write output; returns the number of steps done, or < 0.
 * Energies and observables (capac_md_observables) are those of the last step
 * of the call or of a save step.  Other steps end without waiting for the
 * device, so a RATTLE failure in such a step is reported (as the reference's
 * error) during the next step's force call, before that step's second half,
 * instead of at the end of the failing step; with no CPIM unknowns, or with
 * the per-phase timers on, every step still ends with a synchronisation. */
int capac_md_run(capac_ctx* ctx, int first_step, int n_steps, int n_save, int stop_at_save);

/* Thermostat and step observables: eKsum, instantaneous T, the reference
 * temperature used in the last step, vQ[10], F_max, F_rms (force.c:361-378),
 * pot[15] and vir[9] of the last step (all nullable). */
int capac_md_observables(capac_ctx* ctx, double* eksum, double* temp, double* ref_temp,
                         double* vq, double* fmax, double* frms, double pot[CAPAC_NPOT],
                         double vir[9]);
int capac_set_thermostat(capac_ctx* ctx, const double* vq, double vp);

/* Average device time of one launch of a hot kernel, measured with HIP
 * events on the context's stream over `reps` back-to-back launches on the
 * resident data (no counterpart in the reference; for benchmarks):
 *   CAPAC_KERNEL_SPMV      relay SpMV (one BiCGSTAB half-iteration)
 *   CAPAC_KERNEL_ASSEMBLE  relay assembly
 *   CAPAC_KERNEL_PAIR      pair forces over the resident neighbour rows
 *   CAPAC_KERNEL_NEIGHBOUR cell list + neighbour rows
 *   CAPAC_KERNEL_CPIM_MM   metal-metal CPIM forces
 * Call after one capac_force so that every input is resident. */
#define CAPAC_KERNEL_SPMV      0
#define CAPAC_KERNEL_ASSEMBLE  1
#define CAPAC_KERNEL_PAIR      2
#define CAPAC_KERNEL_NEIGHBOUR 3
#define CAPAC_KERNEL_CPIM_MM   4
#define CAPAC_KERNEL_APPLY64   5   /* FP64 relay product recomputed on the fly (refinement residual) */
#define CAPAC_KERNEL_SOLVE     6   /* the whole CPIM solve of the last force call, nothing else
                                      running: the reference's timer category 7.  After MD steps it
                                      starts from the previous step's solution (the reference's warm
                                      start) whatever the predictor, so both precisions time the
                                      same start; otherwise from the last call's own start */
int capac_time_kernel(capac_ctx* ctx, int kernel, int reps, double* ms_per_launch);

int capac_get_timing(capac_ctx* ctx, capac_timing* t);
int capac_reset_timing(capac_ctx* ctx);
int capac_enable_timing(capac_ctx* ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
