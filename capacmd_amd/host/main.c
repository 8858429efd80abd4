/* capacmd -- CLI of the MI355X build, the reference's program surface
 * (reference mpi-main.c:41-1219, one GPU in place of the MPI ranks):
 *
 *   capacmd -gro init.gro -par param.txt -mds mdset.txt [-prec fp64|fp32] [-dev N]
 *           [-cpt out.cpt] [-rst in.cpt] [-pred 0|1|2]
 *
 * Reads the three inputs with the reference's grammar, runs the force call
 * and then, by run_type, the MD loop (md: traj.gro and vec_pq.txt every 10
 * save steps, traj.dat every save step) or the minimiser (em: steepest
 * descent, cg: conjugate gradients, mpi-main.c:706-865: all three files every
 * step) on the device, and writes the stdout step lines, final energy table
 * and phase timings in the reference's formats.  Failures print the
 * reference's message and exit(1), as the reference does.
 *
 * -pred sets the warm start of the MD solves (capac_set_predictor): 0 the
 * previous solution, as the reference; 1, 2 linear and quadratic
 * extrapolation.  The default is 0 for fp64 and 2 for fp32.
 *
 * Restart (MD only): -cpt writes a checkpoint (capac_checkpoint_save) next
 * to every traj.gro frame and after the last step; -rst continues the run
 * it holds from the step after it, up to number_of_steps, writing the
 * output files from there on.  The continued run repeats the saved run's
 * later steps exactly, which a restart from traj.gro cannot (it keeps vQ[0]
 * and vP of the ten-link chain and no CPIM warm start).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>

#include "capac_gpu.h"
#include "capac_host.h"

static double wall(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

static void die(const char* msg) {
    printf("%s\n", msg);
    exit(1);
}

/* the checkpoint goes to a temporary file that replaces `path` when whole */
static void save_checkpoint(capac_ctx* ctx, const char* path, int step) {
    size_t n = capac_checkpoint_bytes(ctx);
    void* buf = malloc(n);
    if (!buf) die("Error: out of memory for the checkpoint");
    if (capac_checkpoint_save(ctx, step, buf, n) != 0) die(capac_last_error());
    char tmp[4096];
    snprintf(tmp, sizeof tmp, "%s.tmp", path);
    FILE* f = fopen(tmp, "wb");
    if (!f || fwrite(buf, 1, n, f) != n || fclose(f) != 0) {
        printf("Cannot write to %s!\n", path);
        exit(1);
    }
    if (rename(tmp, path) != 0) {
        printf("Cannot write to %s!\n", path);
        exit(1);
    }
    free(buf);
}

static void load_checkpoint(capac_ctx* ctx, const char* path, int* step) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        printf("Cannot read %s!\n", path);
        exit(1);
    }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    void* buf = malloc(n > 0 ? (size_t)n : 1);
    if (!buf || n <= 0 || fread(buf, 1, (size_t)n, f) != (size_t)n) {
        printf("Cannot read %s!\n", path);
        exit(1);
    }
    fclose(f);
    if (capac_checkpoint_load(ctx, buf, (size_t)n, step) != 0) {
        printf("Error: %s: %s\n", path, capac_last_error());
        exit(1);
    }
    free(buf);
}

int main(int argc, char** argv) {
    const char* gro = "init.gro";
    const char* par = "param.txt";
    const char* mds = "mdset.txt";
    int precision = CAPAC_PREC_FP64, device = 0;
    const char* cpt_out = NULL;
    const char* rst_in = NULL;
    int pred = -1;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-gro") && i + 1 < argc) gro = argv[++i];
        else if (!strcmp(argv[i], "-par") && i + 1 < argc) par = argv[++i];
        else if (!strcmp(argv[i], "-mds") && i + 1 < argc) mds = argv[++i];
        else if (!strcmp(argv[i], "-prec") && i + 1 < argc) {
            ++i;
            precision = !strcmp(argv[i], "fp32") ? CAPAC_PREC_FP32_REFINED : CAPAC_PREC_FP64;
        } else if (!strcmp(argv[i], "-dev") && i + 1 < argc) device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-cpt") && i + 1 < argc) cpt_out = argv[++i];
        else if (!strcmp(argv[i], "-rst") && i + 1 < argc) rst_in = argv[++i];
        else if (!strcmp(argv[i], "-pred") && i + 1 < argc) pred = atoi(argv[++i]);
    }
    time_t start_t = time(NULL);
    double t_start = wall();

    char err[1024];
    capac_deck* d = NULL;
    if (capac_deck_load(gro, par, mds, &d, err, sizeof err) != 0) {
        printf("%s", err);
        return 1;
    }
    const capac_params* p = &d->p;
    const int N = p->n_atoms;
    const int n_mat = capac_deck_n_mat(d);

    printf("\n");
    printf("              +-----------------------------------------------------+\n");
    printf("              |      CapacMD force + CPIM path, MI355X (gfx950)     |\n");
    printf("              +-----------------------------------------------------+\n\n");
    printf("    Calculation started at %s", ctime(&start_t));
    printf("    One GPU (device %d), %s relay matrix\n\n", device,
           precision == CAPAC_PREC_FP64 ? "FP64" : "FP32-refined");
    printf("              .------------------ run parameters -------------------.\n\n");
    printf("    run_type = %s, ensemble = %s\n", d->run_type, d->ensemble);
    printf("    vdw_type = %s, coulomb_type = %s\n", d->vdw_type, d->coulomb_type);
    printf("    rCut = %.3f nm, ", p->rcut);
    if (p->use_coulomb == 1) printf("alpha = %.3f nm^-1", p->dsf_alpha);
    printf("\n    ref_T = %.1f K\n\n\n", p->ref_temp);
    printf("              .------------------- molecule info -------------------.\n\n");
    printf("    There are %d types of molecules.\n\n", d->mol_types);
    printf("    %d atoms, %d metal atoms, %d nanoparticles, %d CPIM unknowns\n\n", N, p->n_metal,
           p->n_nps, n_mat);

    const int minimise = !strcmp(d->run_type, "em") || !strcmp(d->run_type, "cg");
    if (strcmp(d->run_type, "md") != 0 && !minimise) {
        printf("Error: unknown run_type %s (md, em or cg)\n", d->run_type);
        return 1;
    }
    if (minimise && (cpt_out || rst_in)) {
        printf("Error: -cpt and -rst continue MD runs (run_type md)\n");
        return 1;
    }

    capac_ctx* ctx = NULL;
    if (capac_create(p, precision, device, &ctx) != 0) {
        printf("Error: %s\n", capac_last_error());
        return 1;
    }
    capac_enable_timing(ctx, 1);
    if (pred >= 0 && capac_set_predictor(ctx, pred) != 0) {
        printf("Error: %s\n", capac_last_error());
        return 1;
    }
    double* x = (double*)calloc(n_mat > 0 ? n_mat : 1, sizeof(double));
    double* r = (double*)malloc(sizeof(double) * 3 * (size_t)N);
    double* v = (double*)malloc(sizeof(double) * 3 * (size_t)N);
    if (capac_set_state(ctx, d->r, d->v, x) != 0) die(capac_last_error());
    double vq[CAPAC_NHC] = {0};
    vq[0] = d->vq0;
    capac_set_thermostat(ctx, vq, d->vp);

    double pot[CAPAC_NPOT], vir[9];
    int rc = 0, first_step = 1;
    if (rst_in) {
        load_checkpoint(ctx, rst_in, &first_step);
        ++first_step;
    } else {
        rc = capac_force(ctx, pot, vir, NULL);
        if (rc == CAPAC_ERR_SOLVER) die(capac_last_error());
        if (rc != 0) { printf("Error: %s\n", capac_last_error()); return 1; }
        if (capac_md_begin(ctx) != 0) die(capac_last_error());
    }

    FILE* fgro = fopen("traj.gro", "w");
    if (!fgro) die("Cannot write to traj.gro!");
    FILE* fpq = fopen("vec_pq.txt", "w");
    if (!fpq) die("Cannot write to vec_pq.txt!");
    FILE* fdat = fopen("traj.dat", "w");
    if (!fdat) die("Cannot write to traj.dat!");

    double ek, temp, reft, fmax, frms;
    capac_md_observables(ctx, &ek, &temp, &reft, vq, &fmax, &frms, pot, vir);
    capac_get_state(ctx, r, v, x);
    if (!rst_in) {
        capac_write_gro(fgro, d, r, v, vq[0], d->vp, 0);
        capac_write_binary(fdat, d, r, 0);
        if (p->n_metal > 0) capac_write_vec_pq(fpq, d, x, 0);
        if (cpt_out) save_checkpoint(ctx, cpt_out, 0);
    }
    printf("              .---------------- start MD calculation ---------------.\n\n");
    if (minimise) {
        printf("    Step %-5d Fmax=%10.3e  E=%15.8e\n", 0, fmax, pot[0]);
        fflush(stdout);
        if (capac_em_begin(ctx) != 0) die(capac_last_error());
        const int cg = !strcmp(d->run_type, "cg");
        double length = d->em_length, tol = d->em_tol;
        for (int step = 1; step <= d->em_steps; ++step) {
            const double old_pot = pot[0];
            rc = capac_em_step(ctx, length, pot, &fmax, &frms);
            if (rc == CAPAC_ERR_SOLVER || rc == CAPAC_ERR_RATTLE1) die(capac_last_error());
            if (rc != 0) { printf("Error: %s\n", capac_last_error()); return 1; }
            const double delta = pot[0] - old_pot;
            length *= delta <= 0.0 ? 1.2 : 0.2;
            printf("    Step %-5d Fmax=%10.3e  E=%15.8e\n", step, fmax, pot[0]);
            fflush(stdout);
            capac_get_state(ctx, r, v, x);
            capac_write_gro(fgro, d, r, v, 0.0, 0.0, step);       /* vQ = vP = 0 in the minimiser */
            capac_write_binary(fdat, d, r, step);
            if (p->n_metal > 0) capac_write_vec_pq(fpq, d, x, step);
            if (fmax < tol && frms < tol * 0.5 && fabs(delta) < tol * 0.1) {
                printf("\n");
                printf("    F_max   (%13.6e) smaller than %e\n", fmax, tol);
                printf("    F_rms   (%13.6e) smaller than %e\n", frms, tol * 0.5);
                printf("    delta_E (%13.6e) smaller than %e\n", delta, tol * 0.1);
                printf("\n");
                printf("    ===========  Optimization converged ============\n");
                break;
            }
            if (capac_em_direction(ctx, cg) != 0) die(capac_last_error());
        }
    }

    int nsave = d->n_save > 0 ? d->n_save : 1;
    int step = first_step;
    if (!minimise && !rst_in) {
        printf("  %10.3f  Fmax=%.2e  E=%.6e  T=%.3e\n", 0.0, fmax, pot[0], temp);
        fflush(stdout);
    }
    while (!minimise && step <= d->n_steps) {
        int done = capac_md_run(ctx, step, d->n_steps - step + 1, nsave, 1);
        if (done < 0) {
            if (done == CAPAC_ERR_SOLVER || done == CAPAC_ERR_RATTLE1 || done == CAPAC_ERR_RATTLE2)
                die(capac_last_error());
            printf("Error: %s\n", capac_last_error());
            return 1;
        }
        step += done;
        int last = step - 1;
        if (last % nsave == 0) {
            capac_md_observables(ctx, &ek, &temp, &reft, vq, &fmax, &frms, pot, vir);
            capac_get_state(ctx, r, v, x);
            printf("  %10.3f  Fmax=%.2e  E=%.6e  T=%.3e (%.1f)\n", p->dt * last, fmax, pot[0], temp, reft);
            fflush(stdout);
            capac_write_binary(fdat, d, r, last);
            if (last % (nsave * 10) == 0) {
                capac_write_gro(fgro, d, r, v, vq[0], d->vp, last);
                if (p->n_metal > 0) capac_write_vec_pq(fpq, d, x, last);
                if (cpt_out) save_checkpoint(ctx, cpt_out, last);
            }
        }
        if (done == 0) break;
    }
    if (!minimise && cpt_out && step > first_step) save_checkpoint(ctx, cpt_out, step - 1);
    fclose(fdat);
    fclose(fpq);
    fclose(fgro);

    capac_md_observables(ctx, &ek, &temp, &reft, vq, &fmax, &frms, pot, vir);
    printf("\n\n              .------------- final potential energies --------------.\n\n");
    capac_print_potential(stdout, pot);
    time_t end_t = time(NULL);
    printf("    Calculation ended normally at %s", ctime(&end_t));
    printf("    %.3f seconds were used\n\n\n", wall() - t_start);
    capac_timing tm;
    capac_get_timing(ctx, &tm);
    printf("              .------------------- time usage ----------------------.\n\n");
    printf("    Total time used in force: %.3f seconds\n", tm.t[0]);
    printf("    Time used in QSC:         %.3f seconds\n", tm.t[2]);
    printf("    Time used in Nonbonded:   %.3f seconds\n", tm.t[4]);
    printf("    Time used in CPIM-matrix: %.3f seconds\n", tm.t[5]);
    printf("    Time used in CPIM-vector: %.3f seconds\n", tm.t[6]);
    printf("    Time used in CPIM-solve:  %.3f seconds (%lld iterations)\n", tm.t[7], tm.solve_iters);
    printf("    Time used in CPIM-force:  %.3f seconds\n", tm.t[8]);
    printf("    Time used in neighbours:  %.3f seconds\n", tm.t[10]);
    printf("    Time used in integrator:  %.3f seconds\n\n", tm.t[12]);
    capac_destroy(ctx);
    capac_deck_free(d);
    free(x); free(r); free(v);
    return 0;
}
