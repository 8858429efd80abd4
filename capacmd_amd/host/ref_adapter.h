/* ref_adapter.h -- capac_params from the reference program's own structs.
 *
 * For a maintainer who keeps the reference's readers (read_settings,
 * read_param_1/2, read_gro, assign_indices; file.c) and replaces mpi_force
 * (force.h:100) with libcapac_gpu.so: after the reference's set-up
 * (mpi-main.c:99-609, one rank) this flattens its per-molecule-type tables
 * into the per-atom and per-molecule arrays capac_create takes.  Compiled as
 * C++ with the reference's sources (its typedef.h holds std::vector members,
 * typedef.h:285-286); include it after the reference's typedef.h.
 */
#ifndef CAPAC_REF_ADAPTER_H
#define CAPAC_REF_ADAPTER_H

#include "capac_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct capac_ref_params capac_ref_params;

/* The arrays live in the returned holder until capac_ref_params_free.  NULL
 * on a deck the device library cannot take (metal not at atoms 0.., more
 * than 32 atoms in a molecule, ...); the message is written to err. */
capac_ref_params* capac_params_from_reference(const Topol* topol, const Atom_Info* atom_info,
                                              const Mol_Info* mol_info, const RunSet* runset,
                                              const Metal* metal, const System* system, char* err,
                                              int errlen);
const capac_params* capac_ref_params_get(const capac_ref_params* h);
void capac_ref_params_free(capac_ref_params* h);

#ifdef __cplusplus
}
#endif
#endif
