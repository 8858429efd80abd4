/* Input decks and output files of CapacMD (see include/capac_host.h).
 *
 * The grammar follows the reference's readers; the code is this project's
 * own: one pass over param.txt into growable tables, then flattening into
 * the capac_params arrays the device library and the oracle consume.
 */
#include "capac_host.h"

#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#define EV2KJMOL_H 96.4853363969351

typedef struct {
    FILE* fh;
    const char* path;
    char line[CAPAC_STR];
    char word[CAPAC_STR];
    char* err;
    int errlen;
} reader;

static int fail(reader* rd, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    if (rd->err && rd->errlen > 0) vsnprintf(rd->err, rd->errlen, fmt, ap);
    va_end(ap);
    return -1;
}

/* next non-blank line; its first word goes to rd->word (file.c:49-67) */
static int next_line(reader* rd) {
    for (;;) {
        if (!fgets(rd->line, CAPAC_STR, rd->fh))
            return fail(rd, "Error while reading file_par!\n");
        rd->word[0] = '\0';
        sscanf(rd->line, "%255s", rd->word);
        if (rd->word[0]) return 0;
    }
}

/* "<NAME> <int>" section header with an expected name */
static int section(reader* rd, const char* name, int* count) {
    if (next_line(rd)) return -1;
    if (strcmp(rd->word, name) != 0)
        return fail(rd, "Error in %s: expected %s in line\n%s", rd->path, name, rd->line);
    if (sscanf(rd->line, "%*s%d", count) != 1) *count = 0;
    return 0;
}

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "Error: out of host memory\n"); exit(1); }
    return p;
}

/* ------------------------------------------------------------ mdset.txt */

static int read_mdset(capac_deck* d, const char* path, char* err, int errlen) {
    FILE* fh = fopen(path, "r");
    if (!fh) {
        snprintf(err, errlen, "Cannot open %s for input!\n", path);
        return -1;
    }
    char line[CAPAC_STR], key[CAPAC_STR], val[CAPAC_STR];
    capac_params* p = &d->p;
    while (fgets(line, CAPAC_STR, fh)) {
        key[0] = val[0] = '\0';
        if (sscanf(line, "%255s%255s", key, val) < 1) continue;
        if (!strcmp(key, "run_type")) strcpy(d->run_type, val);
        else if (!strcmp(key, "ensemble")) strcpy(d->ensemble, val);
        else if (!strcmp(key, "em_steps")) d->em_steps = atoi(val);
        else if (!strcmp(key, "em_length_in_nm")) d->em_length = atof(val);
        else if (!strcmp(key, "em_tolerance")) d->em_tol = atof(val);
        else if (!strcmp(key, "number_of_steps")) d->n_steps = atoi(val);
        else if (!strcmp(key, "time_step_in_ps")) p->dt = atof(val);
        else if (!strcmp(key, "save_step")) d->n_save = atoi(val);
        else if (!strcmp(key, "vdw_type")) strcpy(d->vdw_type, val);
        else if (!strcmp(key, "coulomb_type")) strcpy(d->coulomb_type, val);
        else if (!strcmp(key, "cut_off_radius")) p->rcut = atof(val);
        else if (!strcmp(key, "coulomb_alpha")) p->dsf_alpha = atof(val);
        else if (!strcmp(key, "heating_steps")) p->n_heating = atoi(val);
        else if (!strcmp(key, "ref_T_in_K")) p->ref_temp = atof(val);
        else if (!strcmp(key, "tau_T_in_ps")) p->tau_t = atof(val);
        else if (!strcmp(key, "fix_metal")) p->fix_metal = atoi(val);
        else if (!strcmp(key, "use_cpff")) p->use_cpff = atoi(val);
        /* V/nm -> kJ/mol/(e nm) (file.c:119-123) */
        else if (!strcmp(key, "external_Ex")) p->efield[0] = atof(val) * EV2KJMOL_H;
        else if (!strcmp(key, "external_Ey")) p->efield[1] = atof(val) * EV2KJMOL_H;
        else if (!strcmp(key, "external_Ez")) p->efield[2] = atof(val) * EV2KJMOL_H;
        else printf("Warning: unkonwn option %s in %s\n", key, path);
    }
    fclose(fh);
    if (!strcmp(d->coulomb_type, "cut_off")) p->use_coulomb = 0;
    else if (!strcmp(d->coulomb_type, "wolf_sum")) p->use_coulomb = 1;
    else { snprintf(err, errlen, "Error: unknown coulomb_type %s!\n", d->coulomb_type); return -1; }
    if (!strcmp(d->vdw_type, "cut_off")) p->use_vdw = 0;
    else if (!strcmp(d->vdw_type, "shifted")) p->use_vdw = 1;
    else { snprintf(err, errlen, "Error: unknown vdw_type %s!\n", d->vdw_type); return -1; }
    p->nvt = !strcmp(d->ensemble, "nvt");
    return 0;
}

/* ------------------------------------------------------------ param.txt */

typedef struct { int i, j; double r; } cstr_t;

/* one bonded term or virtual site of a molecule type, local atom indices */
typedef struct { int a[4]; int funct; double par[6]; } term_t;

typedef struct {
    int n_atoms;
    double* q; double* m; int* type;
    unsigned excl[32];          /* excl[a] bit b: a and b excluded */
    int n_cstr; cstr_t* cstr;
    int n_term[5];              /* CAPAC_TERM_*: bonds (funct 5 bonds only
                                   exclude and are not kept), 1-4 pairs,
                                   angles, dihedrals, virtual sites */
    term_t* term[5];
    int n_vsites;
} moltype;

static void moltype_free(moltype* mt, int n) {
    for (int k = 0; mt && k < n; ++k) {
        free(mt[k].q); free(mt[k].m); free(mt[k].type); free(mt[k].cstr);
        for (int t = 0; t < 5; ++t) free(mt[k].term[t]);
    }
    free(mt);
}

/* periodic dihedral (funct 1, 4, 9) to the Ryckaert-Bellemans coefficients
 * that compute_dihedral_9 passes on (force.c:1236-1347), the same operations
 * on the same values; -2 for an n it rejects, -1 for a phi0 it rejects */
int capac_dihedral_periodic_to_rb(double phi0, double kphi, int n, double c[6]) {
    for (int k = 0; k < 6; ++k) c[k] = 0.0;
    if (n < 1 || n > 4) return -2;
    if (phi0 != 0.0 && phi0 != 180.0) return -1;
    const int flip = phi0 == 180.0;
    if (n == 1) { c[0] = kphi; c[1] = flip ? kphi * (-1.0) : kphi; }
    else if (n == 2) { if (flip) { c[0] = kphi * 2.0; c[2] = kphi * (-2.0); } else c[2] = kphi * 2.0; }
    else if (n == 3) {
        c[0] = kphi;
        if (flip) { c[1] = kphi * 3.0; c[3] = kphi * (-4.0); }
        else { c[1] = kphi * (-3.0); c[3] = kphi * 4.0; }
    } else if (n == 4) {
        if (flip) { c[2] = kphi * 8.0; c[4] = kphi * (-8.0); }
        else { c[0] = kphi * 2.0; c[2] = kphi * (-8.0); c[4] = kphi * 8.0; }
    } else return -1;
    c[1] *= -1.0;
    c[3] *= -1.0;
    c[5] *= -1.0;
    return 0;
}

static void excl_pair(moltype* mt, int a, int b) {
    if (a >= 0 && b >= 0 && a < 32 && b < 32) {
        mt->excl[a] |= 1u << b;
        mt->excl[b] |= 1u << a;
    }
}

static int read_param(capac_deck* d, const char* path, moltype** out_mt, char* err, int errlen) {
    reader rd = {0};
    rd.path = path; rd.err = err; rd.errlen = errlen;
    rd.fh = fopen(path, "r");
    if (!rd.fh) { snprintf(err, errlen, "Error: Cannot open param file %s!\n", path); return -1; }
    capac_params* p = &d->p;
    int rc = -1, n;

    if (next_line(&rd)) goto done;
    if (strcmp(rd.word, "SCALE_LJ_QQ")) {
        fail(&rd, "Error in %s: expected SCALE_LJ_QQ in line\n%s", path, rd.line); goto done;
    }
    sscanf(rd.line, "%*s%lf%lf", &d->scale_lj, &d->scale_qq);
    if (section(&rd, "MOLECULES", &d->mol_types)) goto done;
    d->mol_name = xcalloc(d->mol_types, sizeof *d->mol_name);
    d->mol_count = xcalloc(d->mol_types, sizeof(int));
    d->mol_atoms = xcalloc(d->mol_types, sizeof(int));
    moltype* mt = xcalloc(d->mol_types, sizeof(moltype));
    *out_mt = mt;
    for (int k = 0; k < d->mol_types; ++k) {
        if (next_line(&rd)) goto done;
        sscanf(rd.line, "%255s%d", d->mol_name[k], &d->mol_count[k]);
    }
    for (int k = 0; k < d->mol_types; ++k) {
        char name[CAPAC_STR + 16];
        snprintf(name, sizeof name, "%s_ATOMS", d->mol_name[k]);
        if (section(&rd, name, &n)) goto done;
        if (n > 32) { fail(&rd, "Error: molecule %s has %d atoms; at most 32 are supported\n",
                           d->mol_name[k], n); goto done; }
        mt[k].n_atoms = n;
        d->mol_atoms[k] = n;
        mt[k].q = xcalloc(n, sizeof(double));
        mt[k].m = xcalloc(n, sizeof(double));
        mt[k].type = xcalloc(n, sizeof(int));
        for (int a = 0; a < n; ++a) {
            if (next_line(&rd)) goto done;
            sscanf(rd.line, "%lf%lf%d", &mt[k].q[a], &mt[k].m[a], &mt[k].type[a]);
        }
    }
    if (next_line(&rd)) goto done;
    if (strcmp(rd.word, "NONBONDED")) {
        fail(&rd, "Error in %s: expected NONBONDED in line\n%s", path, rd.line); goto done;
    }
    sscanf(rd.line, "%*s%d", &p->n_types);
    {
        int nt = p->n_types;
        d->nb_funct = xcalloc((size_t)nt * nt, sizeof(int));
        d->nb_c6 = xcalloc((size_t)nt * nt, sizeof(double));
        d->nb_c12 = xcalloc((size_t)nt * nt, sizeof(double));
        d->nb_a = xcalloc((size_t)nt * nt, sizeof(double));
        d->nb_b = xcalloc((size_t)nt * nt, sizeof(double));
        for (int i = 0; i < nt; ++i)
            for (int j = i; j < nt; ++j) {
                int f = 0;
                double a = 0, b = 0, c = 0;
                if (next_line(&rd)) goto done;
                sscanf(rd.line, "%d%le%le%le", &f, &a, &b, &c);
                int ij = i * nt + j, ji = j * nt + i;
                d->nb_funct[ij] = f;
                if (f == 1) { d->nb_c6[ij] = a; d->nb_c12[ij] = b; }
                else if (f == 2 || f == 3) { d->nb_a[ij] = a; d->nb_b[ij] = b; d->nb_c6[ij] = c; }
                else if (f == 4) { d->nb_c6[ij] = a; d->nb_c12[ij] = b; d->nb_a[ij] = c; }
                else { fail(&rd, "Error: unsupported vdW type: %d in line:\n%s\n", f, rd.line); goto done; }
                d->nb_funct[ji] = f;
                d->nb_c6[ji] = d->nb_c6[ij]; d->nb_c12[ji] = d->nb_c12[ij];
                d->nb_a[ji] = d->nb_a[ij]; d->nb_b[ji] = d->nb_b[ij];
            }
    }
    for (int k = 0; k < d->mol_types; ++k) {
        static const char* secs[6] = {"BONDS", "PAIRS", "ANGLES", "DIHEDRALS", "VSITES", "CSTRS"};
        for (int s = 0; s < 6; ++s) {
            char name[CAPAC_STR + 16];
            snprintf(name, sizeof name, "%s_%s", d->mol_name[k], secs[s]);
            if (section(&rd, name, &n)) goto done;
            if (s == 5) { mt[k].n_cstr = n; mt[k].cstr = xcalloc(n, sizeof(cstr_t)); }
            else mt[k].term[s] = xcalloc(n, sizeof(term_t));
            for (int t = 0; t < n; ++t) {
                int ai = 0, aj = 0, ak = 0, al = 0, f = 0;
                if (next_line(&rd)) goto done;
                term_t tm;
                memset(&tm, 0, sizeof tm);
                if (s == 0) {               /* bonds: i j funct ... (file.c:721-765) */
                    sscanf(rd.line, "%d%d%d", &ai, &aj, &f);
                    if (f == 1) sscanf(rd.line, "%*d%*d%*d%le%le", &tm.par[0], &tm.par[1]);
                    else if (f == 3) sscanf(rd.line, "%*d%*d%*d%le%le%le", &tm.par[0], &tm.par[1], &tm.par[2]);
                    else if (f != 5) {
                        fail(&rd, "Error: unsupported bond function type: %d in line:\n%s\n", f, rd.line);
                        goto done;
                    }
                    excl_pair(&mt[k], ai, aj);
                    tm.a[0] = ai; tm.a[1] = aj; tm.funct = f;
                    if (f != 5) mt[k].term[0][mt[k].n_term[0]++] = tm;
                } else if (s == 1) {        /* 1-4 pairs (file.c:771-809) */
                    sscanf(rd.line, "%d%d%d", &ai, &aj, &f);
                    if (f == 2)             /* scaleQQ q_i q_j V (sigma) W (epsilon) */
                        sscanf(rd.line, "%*d%*d%*d%lf%lf%lf%le%le", &tm.par[0], &tm.par[1], &tm.par[2],
                               &tm.par[3], &tm.par[4]);
                    else if (f != 1) {
                        fail(&rd, "Error: unsupported pair function type: %d in line:\n%s\n", f, rd.line);
                        goto done;
                    }
                    excl_pair(&mt[k], ai, aj);
                    tm.a[0] = ai; tm.a[1] = aj; tm.funct = f;
                    mt[k].term[1][mt[k].n_term[1]++] = tm;
                } else if (s == 2) {        /* angles exclude all three pairs (file.c:815-859) */
                    sscanf(rd.line, "%d%d%d%d", &ai, &aj, &ak, &f);
                    if (f == 1) sscanf(rd.line, "%*d%*d%*d%*d%le%le", &tm.par[0], &tm.par[1]);
                    else if (f == 5)
                        sscanf(rd.line, "%*d%*d%*d%*d%le%le%le%le", &tm.par[0], &tm.par[1], &tm.par[2],
                               &tm.par[3]);
                    else {
                        fail(&rd, "Error: unsupported angle function type: %d in line:\n%s\n", f, rd.line);
                        goto done;
                    }
                    excl_pair(&mt[k], ai, aj); excl_pair(&mt[k], ai, ak); excl_pair(&mt[k], aj, ak);
                    tm.a[0] = ai; tm.a[1] = aj; tm.a[2] = ak; tm.funct = f;
                    mt[k].term[2][mt[k].n_term[2]++] = tm;
                } else if (s == 3) {        /* dihedrals exclude all six pairs (file.c:865-925) */
                    sscanf(rd.line, "%d%d%d%d%d", &ai, &aj, &ak, &al, &f);
                    if (f == 3) {
                        sscanf(rd.line, "%*d%*d%*d%*d%*d%lf%lf%lf%lf%lf%lf", &tm.par[0], &tm.par[1],
                               &tm.par[2], &tm.par[3], &tm.par[4], &tm.par[5]);
                    } else if (f == 1 || f == 4 || f == 9) {
                        double phi0 = 0.0, kphi = 0.0;
                        int nn = 0;
                        sscanf(rd.line, "%*d%*d%*d%*d%*d%lf%lf%d", &phi0, &kphi, &nn);
                        int bad = capac_dihedral_periodic_to_rb(phi0, kphi, nn, tm.par);
                        if (bad) {
                            fail(&rd, bad == -2 ? "<> Error: n is not equal to 1, 2, 3 or 4!\n"
                                                  "    phi0=%f, kphi=%f, n=%d\n"
                                                : "<> Error: phi0 is not equal to 0 or pi!\n"
                                                  "    phi0=%f, kphi=%f, n=%d\n",
                                 phi0, kphi, nn);
                            goto done;
                        }
                    } else {
                        fail(&rd, "Error: unsupported dihedral function type: %d in line:\n%s\n", f,
                             rd.line);
                        goto done;
                    }
                    excl_pair(&mt[k], ai, aj); excl_pair(&mt[k], ai, ak); excl_pair(&mt[k], ai, al);
                    excl_pair(&mt[k], aj, ak); excl_pair(&mt[k], aj, al); excl_pair(&mt[k], ak, al);
                    tm.a[0] = ai; tm.a[1] = aj; tm.a[2] = ak; tm.a[3] = al; tm.funct = 3;
                    mt[k].term[3][mt[k].n_term[3]++] = tm;
                } else if (s == 4) {        /* TIP5P-type virtual site s excluded from i, j, k */
                    int as = 0;
                    sscanf(rd.line, "%d%d%d%d%d", &ai, &aj, &ak, &as, &f);
                    if (f != 4) {
                        fail(&rd, "Error: unsupported virtual_sites function type: %d in line:\n%s\n",
                             f, rd.line);
                        goto done;
                    }
                    sscanf(rd.line, "%*d%*d%*d%*d%*d%lf%lf%lf", &tm.par[0], &tm.par[1], &tm.par[2]);
                    excl_pair(&mt[k], as, ai); excl_pair(&mt[k], as, aj); excl_pair(&mt[k], as, ak);
                    tm.a[0] = ai; tm.a[1] = aj; tm.a[2] = ak; tm.a[3] = as; tm.funct = 4;
                    mt[k].term[4][mt[k].n_term[4]++] = tm;
                    mt[k].n_vsites++;
                } else {                    /* constraints i j r (file.c:977-989) */
                    double r = 0;
                    sscanf(rd.line, "%d%d%lf", &ai, &aj, &r);
                    mt[k].cstr[t].i = ai; mt[k].cstr[t].j = aj; mt[k].cstr[t].r = r;
                    excl_pair(&mt[k], ai, aj);
                }
            }
        }
    }
    if (next_line(&rd)) goto done;
    if (strcmp(rd.word, "MM_METAL")) {
        fail(&rd, "Error in %s: expected MM_METAL in line\n%s", path, rd.line); goto done;
    }
    d->metal_min = -1; d->metal_max = -1;
    sscanf(rd.line, "%*s%d%d", &d->metal_min, &d->metal_max);
    if (next_line(&rd)) goto done;
    if (strcmp(rd.word, "QSC_METAL")) {
        fail(&rd, "Error in %s: expected QSC_METAL in line\n%s", path, rd.line); goto done;
    }
    sscanf(rd.line, "%*s%d%d%lf%lf%lf", &p->qsc_n, &p->qsc_m, &p->qsc_eps, &p->qsc_c, &p->qsc_a);
    if (next_line(&rd)) goto done;
    if (strcmp(rd.word, "CPFF_METAL")) {
        fail(&rd, "Error in %s: expected CPFF_METAL in line\n%s", path, rd.line); goto done;
    }
    sscanf(rd.line, "%*s%lf%lf", &p->cpff_polar, &p->cpff_capac);
    if (next_line(&rd)) goto done;
    if (strcmp(rd.word, "NANOPARTICLES")) {
        fail(&rd, "Error in %s: expected NANOPARTICLES in line\n%s", path, rd.line); goto done;
    }
    sscanf(rd.line, "%*s%d", &p->n_nps);
    d->np_first = xcalloc(p->n_nps, sizeof(int));
    d->np_last = xcalloc(p->n_nps, sizeof(int));
    d->np_charge = xcalloc(p->n_nps, sizeof(double));
    for (int k = 0; k < p->n_nps; ++k) {
        if (next_line(&rd)) goto done;
        sscanf(rd.line, "%*s%d%d%lf", &d->np_first[k], &d->np_last[k], &d->np_charge[k]);
    }
    rc = 0;
done:
    fclose(rd.fh);
    return rc;
}

/* ------------------------------------------------------------ init.gro */

static int read_gro(capac_deck* d, const char* path, char* err, int errlen) {
    FILE* fh = fopen(path, "r");
    if (!fh) { snprintf(err, errlen, "Cannot open %s for input!\n", path); return -1; }
    char line[CAPAC_STR], w1[CAPAC_STR], w2[CAPAC_STR], w3[CAPAC_STR];
    int step = 0, n = 0, N = d->p.n_atoms;
    d->vq0 = 0.0; d->vp = 0.0;
    if (fgets(line, CAPAC_STR, fh))
        sscanf(line, "%255s%d%255s%lf%255s%lf", w1, &step, w2, &d->vq0, w3, &d->vp);
    if (fgets(line, CAPAC_STR, fh)) sscanf(line, "%d", &n);
    if (n != N) {
        fclose(fh);
        snprintf(err, errlen, "Error: groNAtoms(%d) not equal to nAtoms(%d)!\n", n, N);
        return -1;
    }
    for (int i = 0; i < N; ++i) {
        if (!fgets(line, CAPAC_STR, fh)) break;
        char field[CAPAC_STR];
        size_t len = strlen(line);
        memcpy(field, line, 5); field[5] = 0;
        d->res_id[i] = 0;
        sscanf(field, "%d", &d->res_id[i]);
        memcpy(field, line + 5, 5); field[5] = 0;
        d->res_name[i][0] = 0;
        sscanf(field, "%7s", d->res_name[i]);
        memcpy(field, line + 10, 5); field[5] = 0;
        d->atom_name[i][0] = 0;
        sscanf(field, "%7s", d->atom_name[i]);
        double x[6] = {0, 0, 0, 0, 0, 0};
        if (len > 20) sscanf(line + 20, "%lf%lf%lf%lf%lf%lf", &x[0], &x[1], &x[2], &x[3], &x[4], &x[5]);
        for (int c = 0; c < 3; ++c) {
            d->r[c * N + i] = x[c];
            d->v[c * N + i] = x[3 + c];
        }
    }
    if (fgets(line, CAPAC_STR, fh))
        sscanf(line, "%lf%lf%lf", &d->p.box[0], &d->p.box[1], &d->p.box[2]);
    fclose(fh);
    return 0;
}

/* ------------------------------------------------------------ assembly */

int capac_deck_load(const char* gro, const char* par, const char* mds, capac_deck** out,
                    char* err, int errlen) {
    char dummy[8];
    if (!err) { err = dummy; errlen = sizeof dummy; }
    err[0] = 0;
    *out = NULL;
    capac_deck* d = xcalloc(1, sizeof(capac_deck));
    capac_params* p = &d->p;
    moltype* mt = NULL;
    int rc = read_mdset(d, mds, err, errlen);
    if (!rc) rc = read_param(d, par, &mt, err, errlen);
    if (rc) goto fail_free;

    /* atoms and molecules in param order (file.c:1031-1099) */
    int N = 0, M = 0, C = 0, nvs = 0, nterm[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < d->mol_types; ++k) {
        N += d->mol_count[k] * mt[k].n_atoms;
        M += d->mol_count[k];
        C += d->mol_count[k] * mt[k].n_cstr;
        nvs += d->mol_count[k] * mt[k].n_vsites;
        for (int t = 0; t < 5; ++t) nterm[t] += d->mol_count[k] * mt[k].n_term[t];
    }
    p->n_atoms = N; p->n_mols = M; p->n_cstr = C;
    d->n_vsites = nvs;
    d->charge = xcalloc(N, sizeof(double));
    d->mass = xcalloc(N, sizeof(double));
    d->atomtype = xcalloc(N, sizeof(int));
    d->mol_id = xcalloc(N, sizeof(int));
    d->local_idx = xcalloc(N, sizeof(int));
    d->excl_mask = xcalloc(N, sizeof(unsigned));
    d->mol_first = xcalloc(M, sizeof(int));
    d->mol_last = xcalloc(M, sizeof(int));
    d->mol_cstr = xcalloc(M + 1, sizeof(int));
    d->cstr_i = xcalloc(C, sizeof(int));
    d->cstr_j = xcalloc(C, sizeof(int));
    d->cstr_r = xcalloc(C, sizeof(double));
    d->r = xcalloc(3 * (size_t)N, sizeof(double));
    d->v = xcalloc(3 * (size_t)N, sizeof(double));
    d->res_id = xcalloc(N, sizeof(int));
    d->res_name = xcalloc(N, sizeof *d->res_name);
    d->atom_name = xcalloc(N, sizeof *d->atom_name);
    {
        int i = 0, im = 0, ic = 0;
        for (int k = 0; k < d->mol_types; ++k)
            for (int c = 0; c < d->mol_count[k]; ++c, ++im) {
                d->mol_first[im] = i;
                d->mol_cstr[im] = ic;
                for (int t = 0; t < mt[k].n_cstr; ++t, ++ic) {
                    d->cstr_i[ic] = i + mt[k].cstr[t].i;
                    d->cstr_j[ic] = i + mt[k].cstr[t].j;
                    d->cstr_r[ic] = mt[k].cstr[t].r;
                }
                for (int a = 0; a < mt[k].n_atoms; ++a, ++i) {
                    d->charge[i] = mt[k].q[a];
                    d->mass[i] = mt[k].m[a];
                    d->atomtype[i] = mt[k].type[a];
                    d->mol_id[i] = im;
                    d->local_idx[i] = a;
                    d->excl_mask[i] = mt[k].excl[a];
                }
                d->mol_last[im] = i - 1;
            }
        d->mol_cstr[M] = ic;
    }
    for (int i = 0; i < N; ++i)
        if (d->atomtype[i] < 0 || d->atomtype[i] >= p->n_types) {
            snprintf(err, errlen, "Error: atom %d has atomtype %d outside the %d NONBONDED types\n",
                     i, d->atomtype[i], p->n_types);
            rc = -1; goto fail_free;
        }
    if (d->metal_min >= 0 && d->metal_max >= d->metal_min) {
        if (d->metal_min != 0) {
            snprintf(err, errlen, "Error: metal atoms must be the first atoms (MM_METAL starts at %d)\n",
                     d->metal_min);
            rc = -1; goto fail_free;
        }
        if (d->metal_max >= N) {
            snprintf(err, errlen, "Error: MM_METAL range ends at %d but the system has %d atoms\n",
                     d->metal_max, N);
            rc = -1; goto fail_free;
        }
        p->n_metal = d->metal_max - d->metal_min + 1;
    } else {
        p->n_metal = 0;
    }
    for (int k = 0; k < p->n_nps; ++k)
        if (d->np_first[k] < 0 || d->np_last[k] >= p->n_metal || d->np_first[k] > d->np_last[k]) {
            snprintf(err, errlen, "Error: nanoparticle %d range %d..%d lies outside the metal atoms\n",
                     k, d->np_first[k], d->np_last[k]);
            rc = -1; goto fail_free;
        }
    if (nterm[0] + nterm[1] + nterm[2] + nterm[3] + nterm[4] > 0) {
        /* expand the molecule types' terms to every molecule (mpi_bonded
           and build_vsites walk molecules in order, force.c:1582-1904) */
        d->mol_terms = xcalloc(5 * ((size_t)M + 1), sizeof(int));
        d->bond_atoms = xcalloc(2 * (size_t)nterm[0], sizeof(int));
        d->bond_funct = xcalloc(nterm[0], sizeof(int));
        d->bond_par = xcalloc(3 * (size_t)nterm[0], sizeof(double));
        d->pair14_atoms = xcalloc(2 * (size_t)nterm[1], sizeof(int));
        d->pair14_par = xcalloc(6 * (size_t)nterm[1], sizeof(double));
        d->angle_atoms = xcalloc(3 * (size_t)nterm[2], sizeof(int));
        d->angle_funct = xcalloc(nterm[2], sizeof(int));
        d->angle_par = xcalloc(4 * (size_t)nterm[2], sizeof(double));
        d->dihedral_atoms = xcalloc(4 * (size_t)nterm[3], sizeof(int));
        d->dihedral_par = xcalloc(6 * (size_t)nterm[3], sizeof(double));
        d->vsite_atoms = xcalloc(4 * (size_t)nterm[4], sizeof(int));
        d->vsite_par = xcalloc(3 * (size_t)nterm[4], sizeof(double));
        int at[5] = {0, 0, 0, 0, 0}, im = 0;
        const int nt = p->n_types;
        for (int k = 0; k < d->mol_types; ++k)
            for (int c = 0; c < d->mol_count[k]; ++c, ++im) {
                const int f0 = d->mol_first[im];
                for (int t = 0; t < 5; ++t) d->mol_terms[t * ((size_t)M + 1) + im] = at[t];
                for (int q = 0; q < mt[k].n_term[0]; ++q, ++at[0]) {
                    const term_t* tm = &mt[k].term[0][q];
                    d->bond_atoms[2 * at[0]] = f0 + tm->a[0];
                    d->bond_atoms[2 * at[0] + 1] = f0 + tm->a[1];
                    d->bond_funct[at[0]] = tm->funct;
                    for (int u = 0; u < 3; ++u) d->bond_par[3 * at[0] + u] = tm->par[u];
                }
                for (int q = 0; q < mt[k].n_term[1]; ++q, ++at[1]) {
                    const term_t* tm = &mt[k].term[1][q];
                    const int i = f0 + tm->a[0], j = f0 + tm->a[1];
                    double* pp = d->pair14_par + 6 * at[1];
                    d->pair14_atoms[2 * at[1]] = i;
                    d->pair14_atoms[2 * at[1] + 1] = j;
                    if (tm->funct == 1) {   /* force.c:1752-1765 */
                        const int ij = d->atomtype[i] * nt + d->atomtype[j];
                        pp[0] = d->nb_c6[ij]; pp[1] = d->nb_c12[ij];
                        pp[2] = d->charge[i]; pp[3] = d->charge[j];
                        pp[4] = d->scale_lj; pp[5] = d->scale_qq;
                    } else {                /* user-defined pair (force.c:1768-1782) */
                        const double sig_6 = pow(tm->par[3], 6.0);
                        const double eps_4 = 4.0 * tm->par[4];
                        pp[0] = eps_4 * sig_6; pp[1] = pp[0] * sig_6;
                        pp[2] = tm->par[1]; pp[3] = tm->par[2];
                        pp[4] = 1.0; pp[5] = tm->par[0];
                    }
                }
                for (int q = 0; q < mt[k].n_term[2]; ++q, ++at[2]) {
                    const term_t* tm = &mt[k].term[2][q];
                    for (int u = 0; u < 3; ++u) d->angle_atoms[3 * at[2] + u] = f0 + tm->a[u];
                    d->angle_funct[at[2]] = tm->funct;
                    for (int u = 0; u < 4; ++u) d->angle_par[4 * at[2] + u] = tm->par[u];
                }
                for (int q = 0; q < mt[k].n_term[3]; ++q, ++at[3]) {
                    const term_t* tm = &mt[k].term[3][q];
                    for (int u = 0; u < 4; ++u) d->dihedral_atoms[4 * at[3] + u] = f0 + tm->a[u];
                    for (int u = 0; u < 6; ++u) d->dihedral_par[6 * at[3] + u] = tm->par[u];
                }
                for (int q = 0; q < mt[k].n_term[4]; ++q, ++at[4]) {
                    const term_t* tm = &mt[k].term[4][q];
                    for (int u = 0; u < 4; ++u) d->vsite_atoms[4 * at[4] + u] = f0 + tm->a[u];
                    for (int u = 0; u < 3; ++u) d->vsite_par[3 * at[4] + u] = tm->par[u];
                }
            }
        for (int t = 0; t < 5; ++t) d->mol_terms[t * ((size_t)M + 1) + M] = at[t];
        p->n_bonds = nterm[0]; p->n_pairs14 = nterm[1]; p->n_angles = nterm[2];
        p->n_dihedrals = nterm[3]; p->n_vsites = nterm[4];
        p->mol_terms = d->mol_terms;
        p->bond_atoms = d->bond_atoms; p->bond_funct = d->bond_funct; p->bond_par = d->bond_par;
        p->pair14_atoms = d->pair14_atoms; p->pair14_par = d->pair14_par;
        p->angle_atoms = d->angle_atoms; p->angle_funct = d->angle_funct; p->angle_par = d->angle_par;
        p->dihedral_atoms = d->dihedral_atoms; p->dihedral_par = d->dihedral_par;
        p->vsite_atoms = d->vsite_atoms; p->vsite_par = d->vsite_par;
    }
    p->ndf = 3 * (N - nvs) - 3 - C;            /* mpi-main.c:548 */
    rc = read_gro(d, gro, err, errlen);
    if (rc) goto fail_free;

    p->charge = d->charge; p->mass = d->mass; p->atomtype = d->atomtype;
    p->mol_id = d->mol_id; p->local_idx = d->local_idx; p->excl_mask = d->excl_mask;
    p->mol_first = d->mol_first; p->mol_last = d->mol_last; p->mol_cstr = d->mol_cstr;
    p->cstr_i = d->cstr_i; p->cstr_j = d->cstr_j; p->cstr_r = d->cstr_r;
    p->np_first = d->np_first; p->np_last = d->np_last; p->np_charge = d->np_charge;
    p->nb_funct = d->nb_funct; p->nb_c6 = d->nb_c6; p->nb_c12 = d->nb_c12;
    p->nb_a = d->nb_a; p->nb_b = d->nb_b;
    moltype_free(mt, d->mol_types);
    *out = d;
    return 0;
fail_free:
    moltype_free(mt, d ? d->mol_types : 0);
    capac_deck_free(d);
    return rc ? rc : -1;
}

void capac_deck_free(capac_deck* d) {
    if (!d) return;
    free(d->mol_name); free(d->mol_count); free(d->mol_atoms);
    free(d->charge); free(d->mass); free(d->atomtype); free(d->mol_id); free(d->local_idx);
    free(d->excl_mask); free(d->mol_first); free(d->mol_last); free(d->mol_cstr);
    free(d->cstr_i); free(d->cstr_j); free(d->cstr_r);
    free(d->np_first); free(d->np_last); free(d->np_charge);
    free(d->nb_funct); free(d->nb_c6); free(d->nb_c12); free(d->nb_a); free(d->nb_b);
    free(d->r); free(d->v); free(d->res_id); free(d->res_name); free(d->atom_name);
    free(d->mol_terms); free(d->bond_atoms); free(d->bond_funct); free(d->bond_par);
    free(d->pair14_atoms); free(d->pair14_par); free(d->angle_atoms); free(d->angle_funct);
    free(d->angle_par); free(d->dihedral_atoms); free(d->dihedral_par); free(d->vsite_atoms);
    free(d->vsite_par);
    free(d);
}

int capac_deck_n_mat(const capac_deck* d) {
    return d->p.n_metal > 0 ? 4 * d->p.n_metal + d->p.n_nps : 0;
}

const capac_params* capac_deck_params(const capac_deck* d) { return &d->p; }
const double* capac_deck_positions(const capac_deck* d) { return d->r; }
const double* capac_deck_velocities(const capac_deck* d) { return d->v; }
void capac_deck_run(const capac_deck* d, int* n_steps, int* n_save, double* vq0, double* vp) {
    if (n_steps) *n_steps = d->n_steps;
    if (n_save) *n_save = d->n_save;
    if (vq0) *vq0 = d->vq0;
    if (vp) *vp = d->vp;
}
int capac_deck_minimiser(const capac_deck* d, int* em_steps, double* em_length, double* em_tol) {
    if (em_steps) *em_steps = d->em_steps;
    if (em_length) *em_length = d->em_length;
    if (em_tol) *em_tol = d->em_tol;
    if (!strcmp(d->run_type, "md")) return 0;
    if (!strcmp(d->run_type, "em")) return 1;
    if (!strcmp(d->run_type, "cg")) return 2;
    return -1;
}
int capac_params_size(void) { return (int)sizeof(capac_params); }

/* ------------------------------------------------------------ writers */

void capac_write_gro(FILE* fh, const capac_deck* d, const double* r, const double* v,
                     double vq0, double vp, int step) {
    int N = d->p.n_atoms;
    fprintf(fh, "step  %d  vQ  %lf  vP  %lf\n", step, vq0, vp);
    fprintf(fh, "%5d\n", N);
    for (int i = 0; i < N; ++i)
        fprintf(fh, "%5d%-5s%5s%5d%10.5f%10.5f%10.5f%10.6f%10.6f%10.6f\n",
                d->res_id[i], d->res_name[i], d->atom_name[i], i + 1,
                r[i], r[N + i], r[2 * N + i], v[i], v[N + i], v[2 * N + i]);
    fprintf(fh, "%10.5f%10.5f%10.5f\n", d->p.box[0], d->p.box[1], d->p.box[2]);
}

void capac_write_vec_pq(FILE* fh, const capac_deck* d, const double* x, int step) {
    int nm = d->p.n_metal;
    fprintf(fh, "step %d", step);
    fprintf(fh, "  lambda");
    for (int k = 0; k < d->p.n_nps; ++k) fprintf(fh, " %.10f", x[4 * nm + k]);
    fprintf(fh, "\n");
    fprintf(fh, "%5d\n", nm);
    for (int i = 0; i < nm; ++i)
        fprintf(fh, "%15.8f%15.8f%15.8f%15.8f\n", x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * nm + i]);
}

void capac_write_binary(FILE* fh, const capac_deck* d, const double* r, int step) {
    int N = d->p.n_atoms;
    fwrite(&N, sizeof(int), 1, fh);
    fwrite(&step, sizeof(int), 1, fh);
    fwrite(d->p.box, sizeof(double), 3, fh);
    fwrite(r, sizeof(double), 3 * (size_t)N, fh);   /* rx | ry | rz */
}

int capac_write_gro_path(const char* path, const char* mode, const capac_deck* d,
                         const double* r, const double* v, double vq0, double vp, int step) {
    FILE* fh = fopen(path, mode);
    if (!fh) return -1;
    capac_write_gro(fh, d, r, v, vq0, vp, step);
    fclose(fh);
    return 0;
}

int capac_write_vec_pq_path(const char* path, const char* mode, const capac_deck* d,
                            const double* x, int step) {
    FILE* fh = fopen(path, mode);
    if (!fh) return -1;
    capac_write_vec_pq(fh, d, x, step);
    fclose(fh);
    return 0;
}

int capac_write_binary_path(const char* path, const char* mode, const capac_deck* d,
                            const double* r, int step) {
    FILE* fh = fopen(path, mode);
    if (!fh) return -1;
    capac_write_binary(fh, d, r, step);
    fclose(fh);
    return 0;
}

void capac_print_potential(FILE* fh, const double* pot) {
    fprintf(fh, "    Potential energies: \n");
    fprintf(fh, " %15.6e   total energy                         \n", pot[0]);
    fprintf(fh, " %15.6e   metal quantum Sutton-Chen energy     \n", pot[1]);
    fprintf(fh, " %15.6e   non-metal bond stretching energy     \n", pot[2]);
    fprintf(fh, " %15.6e   non-metal angle bending energy       \n", pot[3]);
    fprintf(fh, " %15.6e   non-metal torsional energy           \n", pot[4]);
    fprintf(fh, " %15.6e   Long range Coulomb                   \n", pot[6]);
    fprintf(fh, " %15.6e   Coulomb energy (including 1-4)       \n", pot[7]);
    fprintf(fh, " %15.6e   vdW energy (including 1-4)           \n", pot[8]);
    fprintf(fh, " %15.6e   CPIM metal charge - non-metal charge \n", pot[10]);
    fprintf(fh, " %15.6e   CPIM metal dipole - non-metal charge \n", pot[11]);
    fprintf(fh, " %15.6e   CPIM metal charge - metal charge     \n", pot[12]);
    fprintf(fh, " %15.6e   CPIM metal charge - metal dipole     \n", pot[13]);
    fprintf(fh, " %15.6e   CPIM metal dipole - metal dipole     \n", pot[14]);
    fprintf(fh, "\n");
}
