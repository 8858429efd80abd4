// capac_params from the reference program's structs (see ref_adapter.h).
//
// Reference layout read here (typedef.h:150-401, filled by file.c:286-1099):
//   Atom_Info[i]   charge, mass, atomtype; iMol (molecule type), iAtom (index
//                  in the molecule), molID (molecule index)   (file.c:1031-1099)
//   Mol_Info[m]    mini, maxi: the molecule's first and last atom
//   Topol          per molecule type: exclude[type][a][b] (typedef.h:301),
//                  constraint, bond/pair/angle/dihedral parameters, vsite_4;
//                  nonbonded_param[ti][tj] {funct, C12, C6, A, B}; scaleLJ/QQ
//   Metal          min..max, n_NPs, start_NP/end_NP/cpff_chg, Sutton-Chen and
//                  CPIM constants, use_cpff, fix_pos
//   RunSet         rCut, w_alpha, use_coulomb, use_vdw, dt, ref_temp, tau_temp,
//                  nHeating, ensemble, external_efield (kJ/mol/(e nm), file.c:88-90)
//   System         box, ndf (mpi-main.c:548)
// Every per-type table is expanded to every molecule with absolute atom
// indices, molecules in order, terms in param.txt order (the order
// mpi_bonded, build_vsites and rattle_* walk them, force.c:1582-2090).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "typedef.h"
#include "ref_adapter.h"
#include "capac_host.h"

struct capac_ref_params {
    capac_params p;
    std::vector<double> charge, mass, cstr_r, np_charge, nb_c6, nb_c12, nb_a, nb_b;
    std::vector<int> atomtype, mol_id, local_idx, mol_first, mol_last, mol_cstr, cstr_i, cstr_j;
    std::vector<int> np_first, np_last, nb_funct;
    std::vector<unsigned> excl_mask;
    std::vector<int> mol_terms, bond_atoms, bond_funct, pair14_atoms, angle_atoms, angle_funct;
    std::vector<int> dihedral_atoms, vsite_atoms;
    std::vector<double> bond_par, pair14_par, angle_par, dihedral_par, vsite_par;
};

template <class T>
static const T* ptr(const std::vector<T>& v) { return v.empty() ? NULL : v.data(); }

extern "C" capac_ref_params* capac_params_from_reference(const Topol* t, const Atom_Info* ai,
                                                         const Mol_Info* mi, const RunSet* rs,
                                                         const Metal* m, const System* s, char* err,
                                                         int errlen) {
    capac_ref_params* h = new capac_ref_params();
    capac_params& p = h->p;
    memset(&p, 0, sizeof p);
    const int N = t->n_atoms, M = t->n_mols, nt = t->n_types;
    p.n_atoms = N;
    p.n_mols = M;
    p.n_types = nt;
    p.n_metal = (m->min >= 0 && m->max >= m->min) ? m->max - m->min + 1 : 0;
    if (p.n_metal > 0 && m->min != 0) {
        snprintf(err, errlen, "Error: metal atoms must be the first atoms (MM_METAL starts at %d)\n", m->min);
        delete h;
        return NULL;
    }
    p.n_nps = m->n_NPs;

    // per atom (assign_indices, file.c:1031-1099); the exclusion rows of the
    // atom's molecule type become a bit mask over the other atoms' local index
    h->charge.resize(N); h->mass.resize(N); h->atomtype.resize(N);
    h->mol_id.resize(N); h->local_idx.resize(N); h->excl_mask.resize(N);
    for (int i = 0; i < N; ++i) {
        const int type = ai[i].iMol, a = ai[i].iAtom;
        const int na = t->atom_num[type];
        if (na > 32) {
            snprintf(err, errlen, "Error: molecule type %d has %d atoms; at most 32 are supported\n", type, na);
            delete h;
            return NULL;
        }
        h->charge[i] = ai[i].charge;
        h->mass[i] = ai[i].mass;
        h->atomtype[i] = ai[i].atomtype;
        h->mol_id[i] = ai[i].molID;
        h->local_idx[i] = a;
        unsigned mask = 0;
        for (int b = 0; b < na; ++b)
            if (t->exclude[type][a][b] == 1) mask |= 1u << b;
        h->excl_mask[i] = mask;
    }

    // per molecule, constraints in param.txt order (rattle_1st, force.c:1907-1997)
    h->mol_first.resize(M); h->mol_last.resize(M); h->mol_cstr.resize(M + 1);
    for (int im = 0; im < M; ++im) {
        const int f0 = mi[im].mini, type = ai[f0].iMol;
        h->mol_first[im] = f0;
        h->mol_last[im] = mi[im].maxi;
        h->mol_cstr[im] = (int)h->cstr_i.size();
        for (int k = 0; k < t->n_constraints[type]; ++k) {
            const Constraint& c = t->constraint[type][k];
            h->cstr_i.push_back(f0 + c.atom_i);
            h->cstr_j.push_back(f0 + c.atom_j);
            h->cstr_r.push_back(c.r);
        }
    }
    h->mol_cstr[M] = (int)h->cstr_i.size();
    p.n_cstr = (int)h->cstr_i.size();

    // nanoparticles (file.c:1003-1008)
    for (int k = 0; k < m->n_NPs; ++k) {
        h->np_first.push_back(m->start_NP[k]);
        h->np_last.push_back(m->end_NP[k]);
        h->np_charge.push_back(m->cpff_chg[k]);
    }

    // pair table (file.c:614-695): the fields each funct reads
    h->nb_funct.assign((size_t)nt * nt, 0);
    h->nb_c6.assign((size_t)nt * nt, 0.0); h->nb_c12.assign((size_t)nt * nt, 0.0);
    h->nb_a.assign((size_t)nt * nt, 0.0); h->nb_b.assign((size_t)nt * nt, 0.0);
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
            const NonbondedParam& q = t->nonbonded_param[a][b];
            const size_t ij = (size_t)a * nt + b;
            h->nb_funct[ij] = q.funct;
            if (q.funct == 1) { h->nb_c6[ij] = q.C6; h->nb_c12[ij] = q.C12; }
            else if (q.funct == 2 || q.funct == 3) { h->nb_a[ij] = q.A; h->nb_b[ij] = q.B; h->nb_c6[ij] = q.C6; }
            else if (q.funct == 4) { h->nb_c6[ij] = q.C6; h->nb_c12[ij] = q.C12; h->nb_a[ij] = q.A; }
        }

    // bonded terms and virtual sites, expanded to every molecule
    h->mol_terms.assign(5 * ((size_t)M + 1), 0);
    for (int im = 0; im < M; ++im) {
        const int f0 = mi[im].mini, type = ai[f0].iMol;
        h->mol_terms[0 * (M + 1) + im] = (int)h->bond_funct.size();
        h->mol_terms[1 * (M + 1) + im] = (int)(h->pair14_atoms.size() / 2);
        h->mol_terms[2 * (M + 1) + im] = (int)h->angle_funct.size();
        h->mol_terms[3 * (M + 1) + im] = (int)(h->dihedral_atoms.size() / 4);
        h->mol_terms[4 * (M + 1) + im] = (int)(h->vsite_atoms.size() / 4);
        for (int k = 0; k < t->n_bonds[type]; ++k) {
            const BondParam& q = t->bond_param[type][k];
            // funct 5 bonds only exclude: the reader stores nothing else for
            // them (file.c:752-756), so neither their funct nor their atoms
            // are set; only funct 1 and 3 carry a term
            if (q.funct != 1 && q.funct != 3) continue;
            h->bond_atoms.push_back(f0 + q.atom_i);
            h->bond_atoms.push_back(f0 + q.atom_j);
            h->bond_funct.push_back(q.funct);
            if (q.funct == 1) { h->bond_par.push_back(q.b0); h->bond_par.push_back(q.kb); h->bond_par.push_back(0.0); }
            else { h->bond_par.push_back(q.b0); h->bond_par.push_back(q.D); h->bond_par.push_back(q.beta); }
        }
        for (int k = 0; k < t->n_pairs[type]; ++k) {
            const PairParam& q = t->pair_param[type][k];
            const int i = f0 + q.atom_i, j = f0 + q.atom_j;
            h->pair14_atoms.push_back(i);
            h->pair14_atoms.push_back(j);
            if (q.funct == 1) {             // the pair table and the SCALE_LJ_QQ line (force.c:1752-1765)
                const size_t ij = (size_t)ai[i].atomtype * nt + ai[j].atomtype;
                h->pair14_par.push_back(h->nb_c6[ij]);
                h->pair14_par.push_back(h->nb_c12[ij]);
                h->pair14_par.push_back(ai[i].charge);
                h->pair14_par.push_back(ai[j].charge);
                h->pair14_par.push_back(t->scaleLJ);
                h->pair14_par.push_back(t->scaleQQ);
            } else {                        // its own sigma, epsilon, charges (force.c:1768-1782)
                const double sig_6 = pow(q.V, 6.0);
                const double eps_4 = 4.0 * q.W;
                h->pair14_par.push_back(eps_4 * sig_6);
                h->pair14_par.push_back(eps_4 * sig_6 * sig_6);
                h->pair14_par.push_back(q.q_i);
                h->pair14_par.push_back(q.q_j);
                h->pair14_par.push_back(1.0);
                h->pair14_par.push_back(q.scaleQQ);
            }
        }
        for (int k = 0; k < t->n_angles[type]; ++k) {
            const AngleParam& q = t->angle_param[type][k];
            h->angle_atoms.push_back(f0 + q.atom_i);
            h->angle_atoms.push_back(f0 + q.atom_j);
            h->angle_atoms.push_back(f0 + q.atom_k);
            h->angle_funct.push_back(q.funct);
            h->angle_par.push_back(q.a0);
            h->angle_par.push_back(q.ka);
            h->angle_par.push_back(q.funct == 5 ? q.b0_ub : 0.0);    // funct 1 reads no Urey-Bradley term
            h->angle_par.push_back(q.funct == 5 ? q.kb_ub : 0.0);
        }
        for (int k = 0; k < t->n_dihedrals[type]; ++k) {
            const DihedralParam& q = t->dihedral_param[type][k];
            h->dihedral_atoms.push_back(f0 + q.atom_i);
            h->dihedral_atoms.push_back(f0 + q.atom_j);
            h->dihedral_atoms.push_back(f0 + q.atom_k);
            h->dihedral_atoms.push_back(f0 + q.atom_l);
            double rb[6] = {q.c0, q.c1, q.c2, q.c3, q.c4, q.c5};
            if (q.funct != 3) {
                // periodic: the Ryckaert-Bellemans coefficients compute_dihedral_9
                // forms at every call (force.c:1236-1347)
                if (capac_dihedral_periodic_to_rb(q.phi0, q.kphi, q.n, rb)) {
                    snprintf(err, errlen, "Error: periodic dihedral phi0=%f, kphi=%f, n=%d is not supported\n",
                             q.phi0, q.kphi, q.n);
                    delete h;
                    return NULL;
                }
            }
            for (int u = 0; u < 6; ++u) h->dihedral_par.push_back(rb[u]);
        }
        for (int k = 0; k < t->n_vsites[type]; ++k) {
            const VSite_4& q = t->vsite_4[type][k];
            h->vsite_atoms.push_back(f0 + q.atom_i);
            h->vsite_atoms.push_back(f0 + q.atom_j);
            h->vsite_atoms.push_back(f0 + q.atom_k);
            h->vsite_atoms.push_back(f0 + q.atom_s);
            h->vsite_par.push_back(q.a);
            h->vsite_par.push_back(q.b);
            h->vsite_par.push_back(q.c);
        }
    }
    h->mol_terms[0 * (M + 1) + M] = (int)h->bond_funct.size();
    h->mol_terms[1 * (M + 1) + M] = (int)(h->pair14_atoms.size() / 2);
    h->mol_terms[2 * (M + 1) + M] = (int)h->angle_funct.size();
    h->mol_terms[3 * (M + 1) + M] = (int)(h->dihedral_atoms.size() / 4);
    h->mol_terms[4 * (M + 1) + M] = (int)(h->vsite_atoms.size() / 4);
    p.n_bonds = (int)h->bond_funct.size();
    p.n_pairs14 = (int)(h->pair14_atoms.size() / 2);
    p.n_angles = (int)h->angle_funct.size();
    p.n_dihedrals = (int)(h->dihedral_atoms.size() / 4);
    p.n_vsites = (int)(h->vsite_atoms.size() / 4);
    if (p.n_bonds + p.n_pairs14 + p.n_angles + p.n_dihedrals + p.n_vsites == 0) h->mol_terms.clear();

    // scalars
    for (int d = 0; d < 3; ++d) p.box[d] = s->box[d];
    p.rcut = rs->rCut;
    p.dsf_alpha = rs->w_alpha;
    p.use_coulomb = rs->use_coulomb;
    p.use_vdw = rs->use_vdw;
    p.qsc_n = m->qsc_n; p.qsc_m = m->qsc_m;
    p.qsc_eps = m->qsc_eps; p.qsc_c = m->qsc_c; p.qsc_a = m->qsc_a;
    p.use_cpff = m->use_cpff;
    p.cpff_polar = m->cpff_polar;
    p.cpff_capac = m->cpff_capac;
    for (int d = 0; d < 3; ++d) p.efield[d] = rs->external_efield[d];
    p.dt = rs->dt;
    p.fix_metal = m->fix_pos;
    p.nvt = strcmp(rs->ensemble, "nvt") == 0;
    p.ref_temp = rs->ref_temp;
    p.tau_t = rs->tau_temp;
    p.n_heating = rs->nHeating;
    p.ndf = s->ndf;

    p.charge = ptr(h->charge); p.mass = ptr(h->mass); p.atomtype = ptr(h->atomtype);
    p.mol_id = ptr(h->mol_id); p.local_idx = ptr(h->local_idx); p.excl_mask = ptr(h->excl_mask);
    p.mol_first = ptr(h->mol_first); p.mol_last = ptr(h->mol_last); p.mol_cstr = ptr(h->mol_cstr);
    p.cstr_i = ptr(h->cstr_i); p.cstr_j = ptr(h->cstr_j); p.cstr_r = ptr(h->cstr_r);
    p.np_first = ptr(h->np_first); p.np_last = ptr(h->np_last); p.np_charge = ptr(h->np_charge);
    p.nb_funct = ptr(h->nb_funct); p.nb_c6 = ptr(h->nb_c6); p.nb_c12 = ptr(h->nb_c12);
    p.nb_a = ptr(h->nb_a); p.nb_b = ptr(h->nb_b);
    p.mol_terms = ptr(h->mol_terms);
    p.bond_atoms = ptr(h->bond_atoms); p.bond_funct = ptr(h->bond_funct); p.bond_par = ptr(h->bond_par);
    p.pair14_atoms = ptr(h->pair14_atoms); p.pair14_par = ptr(h->pair14_par);
    p.angle_atoms = ptr(h->angle_atoms); p.angle_funct = ptr(h->angle_funct); p.angle_par = ptr(h->angle_par);
    p.dihedral_atoms = ptr(h->dihedral_atoms); p.dihedral_par = ptr(h->dihedral_par);
    p.vsite_atoms = ptr(h->vsite_atoms); p.vsite_par = ptr(h->vsite_par);
    return h;
}

extern "C" const capac_params* capac_ref_params_get(const capac_ref_params* h) { return h ? &h->p : NULL; }

extern "C" void capac_ref_params_free(capac_ref_params* h) { delete h; }
