// Sutton-Chen density, embedding and force for the metal atoms: the
// reference's mpi_qsc_dens (force.c:385-461) and mpi_qsc_force (force.c:468-542).
//
// Two passes, one wavefront per metal atom over the metal tail of its
// neighbour row (metal pairs inside rc, the reference's `rij2 < rCut2` set):
//   pass 1: rho_i = sum_j (a/r)^m, inv_sqrt_dens_i = 1/sqrt(rho_i)
//   pass 2: f_i -= eps (n (a/r)^n - c m/2 (rho_i^-1/2 + rho_j^-1/2)(a/r)^m)/r^2 r_ij,
//           with the reference's half virial and pot[1] bookkeeping (force on
//           i only, every ordered pair visited once, as the reference does).
// Integer exponents use repeated squaring.
#include "capac_internal.h"
#include "sc_terms.h"

namespace {

__global__ void __launch_bounds__(256)
k_sc_dens(const double* __restrict__ r, BoxConst bc, const int* __restrict__ n_back,
          const int* __restrict__ nbr, int stride, double* __restrict__ isd) {
    sc_dens_block(blockIdx.x, r, bc, n_back, nbr, stride, isd);
}

__global__ void __launch_bounds__(256)
k_sc_force(const double* __restrict__ r, BoxConst bc, const int* __restrict__ n_back,
           const int* __restrict__ nbr, int stride, const double* __restrict__ isd,
           double* __restrict__ f, int f_stride, int add, double* __restrict__ partial) {
    __shared__ double s_red[4][10];
    sc_force_block(blockIdx.x, r, bc, n_back, nbr, stride, isd, f, f_stride, add, partial, s_red);
}

}  // namespace

// adds the Sutton-Chen force to c->f (f_sc NULL) or writes it to f_sc (3 x
// N_m; the metal-metal CPIM launch adds it, so these launches can run while
// the pair launches still write c->f); red[16..26) = pot1, vir[9]
int sc_forces(capac_ctx* c, double* f_sc) {
    if (c->bc.Nm == 0) return 0;
    int nblk = (c->bc.Nm + 3) / 4;
    k_sc_dens<<<nblk, 256, 0, c->stream>>>(c->r, c->bc, c->n_back, c->nbr, c->nb_stride, c->isd);
    k_sc_force<<<nblk, 256, 0, c->stream>>>(c->r, c->bc, c->n_back, c->nbr, c->nb_stride, c->isd,
                                            f_sc ? f_sc : c->f, f_sc ? c->bc.Nm : c->bc.N, f_sc ? 0 : 1,
                                            partial_seg(c, SEG_SC));
    CHECK_HIP(hipGetLastError());
    return reduce_energies(c, SEG_SC, nblk, 10, 16);
}
