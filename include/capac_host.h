/* capac_host.h -- C host library (libcapac_host.so): CapacMD's file surface.
 *
 * Readers for the three input files and writers for the three output files,
 * with the reference's grammar and formats:
 *   mdset.txt   read_settings      (reference file.c:74-130)
 *   param.txt   read_param_1/2     (reference file.c:286-1023) + assign_indices
 *                                  (file.c:1031-1099)
 *   init.gro    read_gro           (reference file.c:138-203)
 *   traj.gro    write_gro          (reference file.c:212-229)
 *   vec_pq.txt  write_vec_pq       (reference file.c:236-260)
 *   traj.dat    write_binary       (reference file.c:269-278)
 * plus the derived constants of the reference's set-up (mpi-main.c:143-203,
 * 300-323, 546-552).  No GPU code; the CLI (capacmd) and the Python harness
 * share it.
 */
#ifndef CAPAC_HOST_H
#define CAPAC_HOST_H

#include <stdio.h>
#include "capac_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CAPAC_STR 256

typedef struct capac_deck {
    capac_params p;             /* points into the arrays below */

    /* mdset.txt */
    char   run_type[CAPAC_STR], ensemble[CAPAC_STR];
    char   vdw_type[CAPAC_STR], coulomb_type[CAPAC_STR];
    int    em_steps;
    double em_length, em_tol;
    int    n_steps, n_save;

    /* init.gro */
    double* r;                  /* 3N SoA                               */
    double* v;                  /* 3N SoA                               */
    int*    res_id;
    char  (*res_name)[8];
    char  (*atom_name)[8];
    double  vq0, vp;            /* thermostat velocities from the title */

    /* param.txt: molecule types */
    int     mol_types;
    char  (*mol_name)[CAPAC_STR];
    int*    mol_count;          /* molecules of each type               */
    int*    mol_atoms;          /* atoms per molecule of each type      */
    double  scale_lj, scale_qq;
    int     n_vsites;

    /* storage behind p */
    double *charge, *mass, *cstr_r, *np_charge;
    int    *atomtype, *mol_id, *local_idx, *mol_first, *mol_last, *mol_cstr;
    int    *cstr_i, *cstr_j, *np_first, *np_last, *nb_funct;
    unsigned *excl_mask;
    double *nb_c6, *nb_c12, *nb_a, *nb_b;
    int    metal_min, metal_max;
    /* bonded terms and virtual sites, expanded (see capac_params) */
    int    *mol_terms, *bond_atoms, *bond_funct, *pair14_atoms, *angle_atoms, *angle_funct;
    int    *dihedral_atoms, *vsite_atoms;
    double *bond_par, *pair14_par, *angle_par, *dihedral_par, *vsite_par;
} capac_deck;

/* Reads the three input files.  On failure returns < 0 and writes the
 * reference's message (e.g. "Error in param.txt: expected MOLECULES in
 * line ...") into err.  Warnings for unknown mdset keys go to stdout as in
 * the reference. */
int  capac_deck_load(const char* gro, const char* par, const char* mds,
                     capac_deck** out, char* err, int errlen);
void capac_deck_free(capac_deck* d);

int  capac_deck_n_mat(const capac_deck* d);    /* 4 N_m + n_NP (mpi-main.c:573) */

/* accessors for callers that do not mirror the struct (ctypes) */
const capac_params* capac_deck_params(const capac_deck* d);
const double* capac_deck_positions(const capac_deck* d);
const double* capac_deck_velocities(const capac_deck* d);
void capac_deck_run(const capac_deck* d, int* n_steps, int* n_save, double* vq0, double* vp);
/* minimiser settings (mdset.txt em_steps, em_length_in_nm, em_tolerance);
 * returns 0 for run_type md, 1 for em, 2 for cg, -1 otherwise */
int  capac_deck_minimiser(const capac_deck* d, int* em_steps, double* em_length, double* em_tol);
int  capac_params_size(void);                  /* sizeof(capac_params), an ABI check */

/* Writers, byte-compatible with the reference's output files. */
void capac_write_gro(FILE* fh, const capac_deck* d, const double* r, const double* v,
                     double vq0, double vp, int step);
void capac_write_vec_pq(FILE* fh, const capac_deck* d, const double* vec_pq, int step);
void capac_write_binary(FILE* fh, const capac_deck* d, const double* r, int step);

/* path-based wrappers for callers without a FILE* (mode "w" or "a") */
int  capac_write_gro_path(const char* path, const char* mode, const capac_deck* d,
                          const double* r, const double* v, double vq0, double vp, int step);
int  capac_write_vec_pq_path(const char* path, const char* mode, const capac_deck* d,
                             const double* vec_pq, int step);
int  capac_write_binary_path(const char* path, const char* mode, const capac_deck* d,
                             const double* r, int step);

/* The reference's energy table (force.c:337-354). */
void capac_print_potential(FILE* fh, const double* pot);

/* A periodic dihedral (funct 1, 4, 9: phi0 in degrees, kphi, n) as the
 * Ryckaert-Bellemans coefficients rb[0..5] that the reference's
 * compute_dihedral_9 forms from it (force.c:1236-1347); capac_params carries
 * every dihedral in that form.  Returns -2 for an n other than 1-4 and -1 for
 * a phi0 other than 0 or 180, which the reference rejects. */
int  capac_dihedral_periodic_to_rb(double phi0, double kphi, int n, double rb[6]);

#ifdef __cplusplus
}
#endif
#endif
