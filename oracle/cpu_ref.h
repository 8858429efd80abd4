/* cpu_ref.h -- the CPU oracle (test infrastructure only).
 *
 * An FP64 restatement of the reference's per-step force + CPIM path and MD
 * step body, in the reference's loop and summation order, used by tests/,
 * bench.py's cpu_baseline leg and __graft_entry__.smoke() as the checker.
 * Nothing in the product links or calls it.  Each function names the
 * reference lines it follows; the parity pins are listed in cpu_ref.c.
 */
#ifndef CPU_REF_H
#define CPU_REF_H

#include "../include/capac_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oref oref;

oref* oref_create(const capac_params* p);
void  oref_destroy(oref* o);
int   oref_n_mat(const oref* o);

/* 3N SoA state and the CPIM vector (nullable) */
void oref_set_state(oref* o, const double* r, const double* v, const double* vec_pq);
void oref_get_state(const oref* o, double* r, double* v, double* vec_pq, double* f);

/* Emulates the reference's 32-bit pair count (mpi-main.c:406, force.c:1421,
 * 1442) when on; off computes every pair. */
void oref_set_int_pair_count(oref* o, int on);
/* fitting parameters (alpha, c, R_q, R_p, eps_SC, c_SC); widths free */
void oref_set_theta(oref* o, const double* theta6);
long oref_reference_pair_end(int n_atoms);

/* phases; f/pot/vir accumulate (callers zero them) */
void oref_nonb(oref* o, const double* r, double* f, double* pot, double* vir);
void oref_qsc(oref* o, const double* r, double* f, double* pot, double* vir, double* inv_sqrt_dens);
/* mpi_bonded (force.c:1694-1904): adds to f, pot[2,3,4,7,8], vir */
void oref_bonded(oref* o, const double* r, double* f, double* pot, double* vir);
/* build_vsites (force.c:1582-1629) on r; mpi_vsites (force.c:1636-1692) on f */
void oref_vsite_build(oref* o, double* r);
void oref_vsite_spread(oref* o, const double* r, double* f);
void oref_vec_ext(oref* o, const double* r, double* b);
long oref_relay(oref* o, const double* r);            /* returns nnz */
void oref_relay_coo(const oref* o, long* row, long* col, double* val, double* diag);
void oref_relay_dense(const oref* o, double* dense);  /* n_mat x n_mat incl. Lagrange rows */
void oref_relay_apply(const oref* o, const double* x, double* y);   /* y = A x, COO order */
int  oref_solve(oref* o, const double* b, double* x, int* iters, double* rTr, double* bTb);
void oref_cpff_force(oref* o, const double* r, const double* x, double* f, double* pot, double* vir);

/* mpi_force on the oracle's own state; results in the state (pot[0] unsummed) */
int  oref_force(oref* o);
void oref_get_force_out(const oref* o, double* f, double* pot, double* vir, double* vec_ext);
void oref_solve_stats(const oref* o, int* iters, double* rTr, double* bTb, int* kind);

/* MD (mpi-main.c:629-640, 874-1013) */
void oref_md_begin(oref* o);
int  oref_md_step(oref* o, int step, int n_save);
void oref_md_observables(const oref* o, double* eksum, double* temp, double* ref_temp,
                         double* vq, double* fmax, double* frms, double* pot, double* vir);
void oref_set_thermostat(oref* o, const double* vq);

/* phase wall times (s), the reference's timer categories */
void oref_timers(const oref* o, double* t15);

/* solver stop test rTr < crit * bTb (reference: 1e-8) */
void oref_set_solve_crit(oref* o, double crit);

/* energy minimisation, run_type em / cg (mpi-main.c:706-865) */
void oref_em_begin(oref* o);
int  oref_em_step(oref* o, double length, double* pot, double* fmax, double* frms);
void oref_em_direction(oref* o, int cg);

#ifdef __cplusplus
}
#endif
#endif
