// Cell-list neighbour build with the reference's exact cut-off test.
//
// The reference has no neighbour list: every pair loop is a brute-force scan
// followed by `rij2 < rCut2` on pbc_12(r_j - r_i) (force.c:1440-1473,
// 401-414; cpff.c:94-120, 581-596).  Here candidates come from a cell grid of
// width >= rc and are then put through that same FP64 test, evaluated with
// the same operations in the same order (no contraction), so the accepted set
// equals the brute-force set exactly: the single-shift distance is never
// shorter than the minimum-image distance, so every accepted pair lies in
// neighbouring cells.
//
// Rows: for atom i, non-metal neighbours fill row i from the front, metal
// neighbours from the back, so metal-only consumers (Sutton-Chen, the water
// side of the CPIM force) read a short tail.  The row order is deterministic:
// stencil cells in fixed order, atoms in index order within a cell.
#include "capac_internal.h"

namespace {

__global__ void k_cell_assign(const double* __restrict__ r, BoxConst bc, int n0, int n1, int n2,
                              int* __restrict__ cell_of, int* __restrict__ cell_count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bc.N) return;
    int n[3] = {n0, n1, n2};
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double x = r[d * bc.N + i];
        double L = bc.box[d];
        double xw = x - L * floor(x / L);
        double t = xw * n[d] / L;
        // clamp before the conversion (also maps a non-finite coordinate to 0)
        if (!(t >= 0.0)) t = 0.0;
        if (t > (double)(n[d] - 1)) t = (double)(n[d] - 1);
        c[d] = (int)t;
    }
    int cell = (c[0] * n1 + c[1]) * n2 + c[2];
    cell_of[i] = cell;
    atomicAdd(&cell_count[cell], 1);
}

// exclusive scan of cell counts; single workgroup, sequential per thread chunk
__global__ void k_cell_scan(const int* __restrict__ count, int n_cell, int* __restrict__ start,
                            int* __restrict__ fill, int* __restrict__ max_cell) {
    __shared__ int s_part[1024];
    int t = threadIdx.x, nt = blockDim.x;
    int per = (n_cell + nt - 1) / nt;
    int lo = t * per, hi = min(lo + per, n_cell);
    int sum = 0, mx = 0;
    for (int k = lo; k < hi; ++k) { sum += count[k]; mx = max(mx, count[k]); }
    s_part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int k = 0; k < nt; ++k) { int v = s_part[k]; s_part[k] = run; run += v; }
        start[n_cell] = run;
    }
    __syncthreads();
    int run = s_part[t];
    for (int k = lo; k < hi; ++k) { start[k] = run; fill[k] = run; run += count[k]; }
    atomicMax(max_cell, mx);
}

__global__ void k_cell_fill(const int* __restrict__ cell_of, int N, int* __restrict__ fill,
                            int* __restrict__ cell_atoms) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int slot = atomicAdd(&fill[cell_of[i]], 1);
    cell_atoms[slot] = i;
}

// order each cell's atoms by index: rank = number of smaller ids in the cell.
// Cells that fit SORT_LDS ids are ranked from LDS, larger ones from global.
#define SORT_LDS 4096
__global__ void k_cell_sort(const int* __restrict__ start, const int* __restrict__ unsorted,
                            int* __restrict__ cell_atoms) {
    __shared__ int s_ids[SORT_LDS];
    int c = blockIdx.x;
    int lo = start[c], n = start[c + 1] - lo;
    const int* src = unsorted + lo;
    if (n <= SORT_LDS) {
        for (int k = threadIdx.x; k < n; k += blockDim.x) s_ids[k] = src[k];
        __syncthreads();
        src = s_ids;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        int v = src[k], rank = 0;
        for (int m = 0; m < n; ++m) rank += src[m] < v;
        cell_atoms[lo + rank] = v;
    }
}

struct Stencil {
    int n;
    int off[27][3];
};

// One workgroup per cell; each thread owns one atom i of the cell and scans
// the stencil cells, staged through LDS in chunks of blockDim atoms.
__global__ void __launch_bounds__(256)
k_nb_build(const double* __restrict__ r, const int* __restrict__ mol_id,
           const int* __restrict__ local_idx, const unsigned* __restrict__ excl,
           const int* __restrict__ cell_start, const int* __restrict__ cell_atoms,
           BoxConst bc, int n0, int n1, int n2, Stencil st, int stride,
           int* __restrict__ n_front, int* __restrict__ n_back, int* __restrict__ nbr,
           int* __restrict__ overflow) {
    __shared__ double s_x[256], s_y[256], s_z[256];
    __shared__ int s_id[256], s_mol[256], s_loc[256];
    const int N = bc.N, Nm = bc.Nm;
    int cell = blockIdx.x;
    int cz = cell % n2, cy = (cell / n2) % n1, cx = cell / (n1 * n2);
    int lo = cell_start[cell], cnt = cell_start[cell + 1] - lo;
    for (int base = 0; base < cnt; base += blockDim.x) {
        int li = base + threadIdx.x;
        bool active = li < cnt;
        int i = active ? cell_atoms[lo + li] : 0;
        double xi = active ? r[i] : 0.0, yi = active ? r[N + i] : 0.0, zi = active ? r[2 * N + i] : 0.0;
        int moli = active ? mol_id[i] : -1;
        unsigned exi = active ? excl[i] : 0u;
        int nf = 0, nb = 0;
        int* row = nbr + (size_t)i * stride;
        for (int s = 0; s < st.n; ++s) {
            int ox = cx + st.off[s][0], oy = cy + st.off[s][1], oz = cz + st.off[s][2];
            ox = (ox + n0) % n0; oy = (oy + n1) % n1; oz = (oz + n2) % n2;
            int oc = (ox * n1 + oy) * n2 + oz;
            int olo = cell_start[oc], ocnt = cell_start[oc + 1] - olo;
            for (int cb = 0; cb < ocnt; cb += 256) {
                __syncthreads();
                int k = cb + threadIdx.x;
                if (k < ocnt) {
                    int j = cell_atoms[olo + k];
                    s_id[threadIdx.x] = j;
                    s_x[threadIdx.x] = r[j];
                    s_y[threadIdx.x] = r[N + j];
                    s_z[threadIdx.x] = r[2 * N + j];
                    s_mol[threadIdx.x] = mol_id[j];
                    s_loc[threadIdx.x] = local_idx[j];
                }
                __syncthreads();
                if (!active) continue;
                int m = min(256, ocnt - cb);
                for (int q = 0; q < m; ++q) {
                    int j = s_id[q];
                    if (j == i) continue;
                    if (s_mol[q] == moli && ((exi >> s_loc[q]) & 1u)) continue;
                    double dx = s_x[q] - xi, dy = s_y[q] - yi, dz = s_z[q] - zi;
                    pbc12(bc, dx, dy, dz);
                    double r2 = dist2_exact(dx, dy, dz);
                    if (!(r2 < bc.rc2)) continue;
                    if (nf + nb >= stride) { overflow[0] = 1; continue; }
                    if (j < Nm) { ++nb; row[stride - nb] = j; }
                    else { row[nf] = j; ++nf; }
                }
            }
        }
        if (active) { n_front[i] = nf; n_back[i] = nb; }
    }
}

}  // namespace

// Cell grid: width >= rc (1 + 1e-9) so that rounding in the binning can never
// separate an in-range pair by more than one cell; along an axis with fewer
// than 3 cells every cell is a neighbour and duplicates are dropped.
static void make_stencil(const int n[3], Stencil* st) {
    int offs[3][3], cnt[3];
    for (int d = 0; d < 3; ++d) {
        if (n[d] >= 3) { offs[d][0] = -1; offs[d][1] = 0; offs[d][2] = 1; cnt[d] = 3; }
        else if (n[d] == 2) { offs[d][0] = 0; offs[d][1] = 1; cnt[d] = 2; }
        else { offs[d][0] = 0; cnt[d] = 1; }
    }
    st->n = 0;
    for (int a = 0; a < cnt[0]; ++a)
        for (int b = 0; b < cnt[1]; ++b)
            for (int c = 0; c < cnt[2]; ++c) {
                st->off[st->n][0] = offs[0][a];
                st->off[st->n][1] = offs[1][b];
                st->off[st->n][2] = offs[2][c];
                st->n++;
            }
}

int nb_grid_setup(capac_ctx* c) {
    for (int d = 0; d < 3; ++d) {
        int n = (int)floor(c->bc.box[d] / (c->bc.rc * (1.0 + 1e-9)));
        c->ncell[d] = n < 1 ? 1 : n;
    }
    c->n_cell = c->ncell[0] * c->ncell[1] * c->ncell[2];
    return 0;
}

int nb_build(capac_ctx* c) {
    const int N = c->bc.N;
    hipStream_t s = c->stream;
    int rc = fill_zero_i(c, c->cell_count, c->n_cell + 1);
    if (!rc) rc = fill_zero_i(c, c->nb_overflow, 2);
    if (rc) return rc;
    int tb = 256, gb = (N + tb - 1) / tb;
    k_cell_assign<<<gb, tb, 0, s>>>(c->r, c->bc, c->ncell[0], c->ncell[1], c->ncell[2], c->cell_of,
                                    c->cell_count);
    k_cell_scan<<<1, 1024, 0, s>>>(c->cell_count, c->n_cell, c->cell_start, c->cell_fill,
                                   c->nb_overflow + 1);
    k_cell_fill<<<gb, tb, 0, s>>>(c->cell_of, N, c->cell_fill, c->cell_tmp);
    k_cell_sort<<<c->n_cell, 256, 0, s>>>(c->cell_start, c->cell_tmp, c->cell_atoms);
    Stencil st;
    make_stencil(c->ncell, &st);
    k_nb_build<<<c->n_cell, 256, 0, s>>>(c->r, c->mol_id, c->local_idx, c->excl, c->cell_start,
                                         c->cell_atoms, c->bc, c->ncell[0], c->ncell[1], c->ncell[2],
                                         st, c->nb_stride, c->n_front, c->n_back, c->nbr,
                                         c->nb_overflow);
    CHECK_HIP(hipGetLastError());
    return 0;
}
