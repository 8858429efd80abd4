// Cell-list neighbour build with the reference's exact cut-off test.
//
// The reference has no neighbour list: every pair loop is a brute-force scan
// followed by `rij2 < rCut2` on pbc_12(r_j - r_i) (force.c:1440-1473,
// 401-414; cpff.c:94-120, 581-596).  Here candidates come from a cell grid
// and are then put through that same FP64 test, evaluated with the same
// operations in the same order (no contraction), so the accepted set equals
// the brute-force set exactly: the single-shift distance is never shorter
// than the minimum-image distance, so every accepted pair lies within the
// stencil.
//
// Verlet skin: rows hold the pairs within rl = rc + skin (the same exact test
// with rl^2), and every force kernel applies the reference's r^2 < rc^2
// itself, so the pairs used are exactly the brute-force set as long as no
// atom has moved more than skin/2 since the build.  With box >= 2 rl the
// single-shift image is then the same at build time and now for every pair
// that can come within rc.  The rebuild decision is made on the device (the
// largest raw displacement against the positions of the last build); the
// build kernels return at once when no rebuild is due, so no host round trip
// is added.  Smaller boxes use no skin (a build at every force call).
//
// Grid: cells of width >= rl (1 + 1e-9) / 2 and a +-2 stencil, so rounding in
// the binning cannot separate an in-range pair by more than two cells; along
// an axis with fewer than 5 cells every cell is visited once.
//
// Build: one workgroup per home cell; the stencil's atoms are staged through
// LDS once and shared by every atom of the cell; each wavefront tests 64
// candidates per pass and compacts the accepted ones with a ballot, so each
// row is written in contiguous runs.
// Rows: non-metal neighbours fill row i from the front, metal neighbours from
// the back, so metal-only consumers (Sutton-Chen, the water side of the CPIM
// force) read a short tail.  The order is deterministic: stencil cells in a
// fixed order, atoms by index within a cell.
#include "capac_internal.h"

namespace {

constexpr int kHalf = 2;                 // stencil half-width in cells
constexpr int kHistLds = 4096;           // cells handled by the LDS histogram

// a coordinate wrapped into [0, L] (L itself possible by rounding)
__device__ __forceinline__ double wrap_coord(double x, double L) { return x - L * floor(x / L); }

__device__ __forceinline__ int cell_of_pos(const BoxConst& bc, double x, double y, double z,
                                           int n0, int n1, int n2) {
    int n[3] = {n0, n1, n2};
    double p[3] = {x, y, z};
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double L = bc.box[d];
        double xw = wrap_coord(p[d], L);
        double t = xw * n[d] / L;
        // clamp before the conversion (also maps a non-finite coordinate to 0)
        if (!(t >= 0.0)) t = 0.0;
        if (t > (double)(n[d] - 1)) t = (double)(n[d] - 1);
        c[d] = (int)t;
    }
    return (c[0] * n1 + c[1]) * n2 + c[2];
}

// cell of every atom; counts through a per-block LDS histogram so the global
// counters see one atomic per (block, occupied cell)
__global__ void __launch_bounds__(256)
k_cell_assign(const double* __restrict__ r, BoxConst bc, int n0, int n1, int n2, int n_cell,
              int* __restrict__ cell_of, int* __restrict__ rank_in_block,
              int* __restrict__ cell_count, const unsigned long long* __restrict__ need) {
    __shared__ int s_hist[kHistLds];
    if (!need[1]) return;
    const int N = bc.N;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool lds = n_cell <= kHistLds;
    if (lds) {
        for (int k = threadIdx.x; k < n_cell; k += blockDim.x) s_hist[k] = 0;
        __syncthreads();
    }
    int cell = -1;
    if (i < N) {
        cell = cell_of_pos(bc, r[i], r[N + i], r[2 * N + i], n0, n1, n2);
        cell_of[i] = cell;
        if (lds) rank_in_block[i] = atomicAdd(&s_hist[cell], 1);
        else rank_in_block[i] = -1;
    }
    if (lds) {
        __syncthreads();
        for (int k = threadIdx.x; k < n_cell; k += blockDim.x)
            if (s_hist[k]) atomicAdd(&cell_count[k], s_hist[k]);
    } else if (i < N) {
        atomicAdd(&cell_count[cell], 1);
    }
}

// exclusive scan of cell counts; single workgroup
__global__ void k_cell_scan(const int* __restrict__ count, int n_cell, int* __restrict__ start,
                            int* __restrict__ fill, const unsigned long long* __restrict__ need) {
    __shared__ int s_part[1024];
    if (!need[1]) return;
    int t = threadIdx.x, nt = blockDim.x;
    int per = (n_cell + nt - 1) / nt;
    int lo = t * per, hi = min(lo + per, n_cell);
    int sum = 0;
    for (int k = lo; k < hi; ++k) sum += count[k];
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
}

// slots: one global atomic per (block, occupied cell) reserves a range, the
// block's atoms take consecutive slots in it by their LDS rank
__global__ void __launch_bounds__(256)
k_cell_fill(const int* __restrict__ cell_of, const int* __restrict__ rank_in_block, int N,
            int n_cell, int* __restrict__ fill, int* __restrict__ cell_tmp,
            const unsigned long long* __restrict__ need) {
    __shared__ int s_hist[kHistLds];
    __shared__ int s_base[kHistLds];
    if (!need[1]) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_cell <= kHistLds) {
        for (int k = threadIdx.x; k < n_cell; k += blockDim.x) s_hist[k] = 0;
        __syncthreads();
        int cell = i < N ? cell_of[i] : -1;
        if (cell >= 0) atomicAdd(&s_hist[cell], 1);
        __syncthreads();
        for (int k = threadIdx.x; k < n_cell; k += blockDim.x)
            if (s_hist[k]) s_base[k] = atomicAdd(&fill[k], s_hist[k]);
        __syncthreads();
        if (cell >= 0) cell_tmp[s_base[cell] + rank_in_block[i]] = i;
    } else if (i < N) {
        int slot = atomicAdd(&fill[cell_of[i]], 1);
        cell_tmp[slot] = i;
    }
}

// order each cell's atoms by index (rank = number of smaller ids in the cell)
// and gather the cell-sorted SoA copies used by the candidate scan
#define SORT_LDS 4096
__global__ void k_cell_sort(const int* __restrict__ start, const int* __restrict__ unsorted,
                            const double* __restrict__ r, const int* __restrict__ mol_id,
                            const int* __restrict__ local_idx, int N, int* __restrict__ cell_atoms,
                            double* __restrict__ sx, double* __restrict__ sy, double* __restrict__ sz,
                            int* __restrict__ smol, int* __restrict__ sloc, double3 box,
                            const unsigned long long* __restrict__ need) {
    __shared__ int s_ids[SORT_LDS];
    if (!need[1]) return;
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
        int slot = lo + rank;
        cell_atoms[slot] = v;
        // coordinates wrapped into the box as the cell assignment wraps
        // them, so a stencil cell's periodic image is one shift for all its atoms
        sx[slot] = wrap_coord(r[v], box[0]);
        sy[slot] = wrap_coord(r[N + v], box[1]);
        sz[slot] = wrap_coord(r[2 * N + v], box[2]);
        smol[slot] = mol_id[v];
        sloc[slot] = local_idx[v];
    }
}

struct Stencil {
    int n;
    short off[125][3];
};

// One workgroup per home cell.  The stencil cells' atoms are staged through
// LDS as one flat, densely packed candidate array (up to kCap at a time, in
// stencil order), so every atom of the home cell reuses each staged
// candidate; each wavefront owns a set of home atoms and tests 64 staged
// candidates per pass, compacting the accepted ones with a ballot.
constexpr int kCap = 1024;        // staged candidates per pass
constexpr int kMaxHome = 512;     // home atoms per pass
// eight wavefronts per cell: the candidate loop is latency-bound, and more
// waves per compute unit hide it (Au561: 383 us per build at four, 259 at
// eight, 280 at sixteen)
constexpr int kBuildThreads = 512;

// a staged candidate, read with two 16-byte LDS loads; ml = molecule * 32 +
// index in the molecule (the exclusion masks are 32-bit)
struct alignas(16) Cand {
    double x, y, z;
    int id, ml;
};

// Shifted: every box edge holds at least 2 kHalf + 1 cells, so each stencil
// cell is a distinct cell and its periodic image relative to the home cell is
// fixed by the offset's wrap; the staged candidates carry that image (wrapped
// coordinate plus the shift), and the candidate loop needs no periodic
// correction.  Membership can differ from the corrected difference only by
// rounding at the list radius, which the rebuild test's 1e-6 margin covers.
// Otherwise (a stencil that lists every cell once) the loop applies pbc12.
template <bool Shifted>
__global__ void __launch_bounds__(kBuildThreads)
k_nb_build(const double* __restrict__ sx, const double* __restrict__ sy,
           const double* __restrict__ sz, const int* __restrict__ smol,
           const int* __restrict__ sloc, const int* __restrict__ cell_atoms,
           const int* __restrict__ cell_start, const unsigned* __restrict__ excl, BoxConst bc,
           int n0, int n1, int n2, const Stencil* __restrict__ st, int stride,
           int* __restrict__ n_front, int* __restrict__ n_back, int* __restrict__ nbr,
           int* __restrict__ overflow, int* __restrict__ tail_list,
           const unsigned long long* __restrict__ need) {
    __shared__ Cand c_c[kCap];
    if (!need[1]) return;
    __shared__ int h_nf[kMaxHome], h_nb[kMaxHome];
    __shared__ int seg_lo[125], seg_off[126];
    __shared__ double seg_sh[125][3];
    const int Nm = bc.Nm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int cell = blockIdx.x;
    const int cz = cell % n2, cy = (cell / n2) % n1, cx = cell / (n1 * n2);
    const int hlo = cell_start[cell], hhi = cell_start[cell + 1];
    if (hlo == hhi) return;
    const int ns = st->n;
    // stencil cells as segments of one flat candidate index space, in order
    if (threadIdx.x < ns) {
        int t = threadIdx.x;
        int ox = cx + st->off[t][0], oy = cy + st->off[t][1], oz = cz + st->off[t][2];
        if (Shifted) {
            seg_sh[t][0] = ox < 0 ? -bc.box[0] : (ox >= n0 ? bc.box[0] : 0.0);
            seg_sh[t][1] = oy < 0 ? -bc.box[1] : (oy >= n1 ? bc.box[1] : 0.0);
            seg_sh[t][2] = oz < 0 ? -bc.box[2] : (oz >= n2 ? bc.box[2] : 0.0);
        }
        ox = (ox % n0 + n0) % n0; oy = (oy % n1 + n1) % n1; oz = (oz % n2 + n2) % n2;
        const int oc = (ox * n1 + oy) * n2 + oz;
        seg_lo[t] = cell_start[oc];
        seg_off[t + 1] = cell_start[oc + 1] - cell_start[oc];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        seg_off[0] = 0;
        for (int t = 0; t < ns; ++t) seg_off[t + 1] += seg_off[t];
    }
    __syncthreads();
    const int total_all = seg_off[ns];
    for (int hbase = hlo; hbase < hhi; hbase += kMaxHome) {
        const int nh = min(kMaxHome, hhi - hbase);
        for (int h = threadIdx.x; h < nh; h += blockDim.x) { h_nf[h] = 0; h_nb[h] = 0; }
        for (int w0 = 0; w0 < total_all; w0 += kCap) {
            const int total = min(kCap, total_all - w0);
            // padded to whole passes of 64 with far-away entries that no test
            // accepts, so the candidate loop needs no bounds test
            const int padded = (total + 63) & ~63;
            __syncthreads();
            for (int f = total + threadIdx.x; f < padded; f += blockDim.x)
                c_c[f] = Cand{1e30, 1e30, 1e30, -1, -1};
            for (int f = threadIdx.x; f < total; f += blockDim.x) {
                const int g = w0 + f;
                int lo = 0, hi = ns - 1;             // segment holding flat index g
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (seg_off[mid] <= g) lo = mid; else hi = mid - 1;
                }
                const int k = seg_lo[lo] + (g - seg_off[lo]);
                if (Shifted)
                    c_c[f] = Cand{sx[k] + seg_sh[lo][0], sy[k] + seg_sh[lo][1], sz[k] + seg_sh[lo][2],
                                  cell_atoms[k], smol[k] * 32 + sloc[k]};
                else
                    c_c[f] = Cand{sx[k], sy[k], sz[k], cell_atoms[k], smol[k] * 32 + sloc[k]};
            }
            __syncthreads();
            // the home atom is the same on every lane: its index, data and
            // row counts live in scalar registers, so the row-capacity test
            // below is a scalar branch
            for (int h = __builtin_amdgcn_readfirstlane(wave); h < nh; h += kBuildThreads / 64) {
                const int hs = hbase + h;
                const int i = cell_atoms[hs];
                const double xi = sx[hs], yi = sy[hs], zi = sz[hs];
                const int moli = smol[hs];
                const unsigned exi = excl[i];
                int* row = nbr + (size_t)i * stride;
                int nf = __builtin_amdgcn_readfirstlane(h_nf[h]), nb = __builtin_amdgcn_readfirstlane(h_nb[h]);
                if (nf < 0) continue;                 // this row already overflowed
                for (int base = 0; base < padded; base += 64) {
                    // every lane evaluates its candidate without branching;
                    // the tests combine into one predicate
                    const Cand cd = c_c[base + lane];
                    const int j = cd.id;
                    const bool excluded = ((cd.ml >> 5) == moli) && ((exi >> (cd.ml & 31)) & 1u);
                    double dx = cd.x - xi, dy = cd.y - yi, dz = cd.z - zi;
                    if (!Shifted) pbc12(bc, dx, dy, dz);
                    const double r2 = dist2_exact(dx, dy, dz);
                    const bool acc = (j != i) & !excluded & (r2 < bc.rl2);
                    const unsigned long long mf = __ballot(acc && j >= Nm);
                    const unsigned long long mb = __ballot(acc && j < Nm);
                    const int cf = __popcll(mf), cb = __popcll(mb);
                    if (nf + nb + cf + cb > stride) {
                        if (lane == 0) overflow[0] = 1;
                        nf = -1;
                        break;
                    }
                    const int pf = nf + __popcll(mf & below);
                    const int pb = stride - 1 - (nb + __popcll(mb & below));
                    if (acc) row[j >= Nm ? pf : pb] = j;
                    nf += cf;
                    nb += cb;
                }
                if (lane == 0) { h_nf[h] = nf; h_nb[h] = nb; }
            }
        }
        __syncthreads();
        for (int h = threadIdx.x; h < nh; h += blockDim.x) {
            const int i = cell_atoms[hbase + h];
            const int nbv = h_nf[h] < 0 ? 0 : h_nb[h];
            n_front[i] = h_nf[h] < 0 ? 0 : h_nf[h];
            n_back[i] = nbv;
            // non-metal atoms with metal neighbours: the atom-side CPIM
            // force kernel visits only these (their order does not matter,
            // each is summed on its own)
            if (i >= Nm && nbv > 0) tail_list[1 + atomicAdd(&tail_list[0], 1)] = i;
        }
        __syncthreads();
    }
}

// largest squared raw displacement since the last build (atomicMax on the
// bit pattern: non-negative doubles order like their bits) and, in the last
// workgroup of the launch to finish (a completion count in ctl[3]), the
// rebuild decision and the clearing of the counters the build accumulates
// into: one launch where a displacement pass and a decision kernel were two
constexpr int kDispThreads = 1024;
__global__ void __launch_bounds__(kDispThreads)
k_disp_decide(const double* __restrict__ r, const double* __restrict__ r_build, int N,
              unsigned long long* __restrict__ ctl, double limit2, int* __restrict__ cell_count,
              int n_cell, int* __restrict__ overflow, int* __restrict__ tail_list) {
    __shared__ int s_flag;
    __shared__ double s_max[kDispThreads / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (i < N) {
        double dx = r[i] - r_build[i], dy = r[N + i] - r_build[N + i], dz = r[2 * N + i] - r_build[2 * N + i];
        d2 = dx * dx + dy * dy + dz * dz;
        if (!(d2 == d2)) d2 = 1e300;           // non-finite positions force a rebuild
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = d2;
    __syncthreads();
    // one atomic and one fence per workgroup: the global maximum and the
    // completion count are contended by every workgroup of the launch
    if (threadIdx.x == 0) {
        for (int w = 1; w < kDispThreads / 64; ++w) d2 = fmax(d2, s_max[w]);
        atomicMax(&ctl[0], (unsigned long long)__double_as_longlong(d2));
        __threadfence();
        s_flag = atomicAdd(&ctl[3], 1ull) == (unsigned long long)(gridDim.x - 1);
    }
    __syncthreads();
    if (!s_flag) return;
    // the last workgroup: every other one has added its maximum
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const double m2 = __longlong_as_double((long long)atomicMax(&ctl[0], 0ull));
        const int go = ctl[1] != 0ull || m2 > limit2;
        if (go) ctl[1] = 1ull;
        ctl[0] = 0ull;
        ctl[3] = 0ull;
        s_flag = go;
    }
    __syncthreads();
    if (!s_flag) return;
    for (int k = threadIdx.x; k <= n_cell; k += blockDim.x) cell_count[k] = 0;
    if (threadIdx.x < 2) overflow[threadIdx.x] = 0;
    if (threadIdx.x == 0) tail_list[0] = 0;
}

// rebuild flag from the displacement; when set, also clears the counters the
// build accumulates into
__global__ void k_decide(unsigned long long* __restrict__ ctl, double limit2, int* __restrict__ cell_count,
                         int n_cell, int* __restrict__ overflow, int* __restrict__ tail_list) {
    __shared__ int s_go;
    if (threadIdx.x == 0) {
        double d2 = __longlong_as_double((long long)ctl[0]);
        int go = ctl[1] != 0ull || d2 > limit2;
        if (go) ctl[1] = 1ull;
        ctl[0] = 0ull;
        s_go = go;
    }
    __syncthreads();
    if (!s_go) return;
    for (int k = threadIdx.x; k <= n_cell; k += blockDim.x) cell_count[k] = 0;
    if (threadIdx.x < 2) overflow[threadIdx.x] = 0;
    if (threadIdx.x == 0) tail_list[0] = 0;
}

// after a build: remember the positions; the last workgroup to finish (a
// completion count in ctl[4]) clears the flag, once every workgroup has read it
__global__ void __launch_bounds__(256)
k_mark_built(const double* __restrict__ r, double* __restrict__ r_build, int n3,
             unsigned long long* __restrict__ ctl) {
    __shared__ int s_flag;
    if (threadIdx.x == 0) s_flag = ctl[1] != 0ull;
    __syncthreads();
    if (!s_flag) return;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n3) r_build[k] = r[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctl[4], 1ull) == (unsigned long long)(gridDim.x - 1)) {
            ctl[4] = 0ull;
            ctl[1] = 0ull;
            ctl[2] += 1ull;
        }
    }
}

__global__ void k_set_flag(unsigned long long* __restrict__ ctl) {
    if (threadIdx.x == 0) { ctl[0] = 0ull; ctl[1] = 1ull; }
}

}  // namespace

static void make_stencil(const int n[3], Stencil* st) {
    short offs[3][5];
    int cnt[3];
    for (int d = 0; d < 3; ++d) {
        if (n[d] >= 2 * kHalf + 1) {
            cnt[d] = 2 * kHalf + 1;
            for (int k = 0; k < cnt[d]; ++k) offs[d][k] = (short)(k - kHalf);
        } else {
            cnt[d] = n[d];                       // every cell once
            for (int k = 0; k < cnt[d]; ++k) offs[d][k] = (short)k;
        }
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

// list radius: rc + skin where every box edge is at least 2 (rc + skin), else rc
int nb_set_radius(capac_ctx* c, double skin) {
    double rl = c->bc.rc + skin;
    for (int d = 0; d < 3; ++d)
        if (c->bc.box[d] < 2.0 * rl) { skin = 0.0; rl = c->bc.rc; }
    c->skin = skin;
    // Without a skin the rows hold the pairs within rc itself.  The build's
    // distance test works on wrapped and shifted coordinates, which round
    // differently from the force kernels' pbc_12 differences, so its radius
    // is rc widened by a relative 1e-12: a pair within rounding of rc stays
    // in the row, and the force kernels apply the exact r^2 < rc^2 test.
    c->bc.rl2 = skin > 0.0 ? rl * rl : c->bc.rc2 * (1.0 + 1e-12);
    return nb_grid_setup(c);
}

int nb_grid_setup(capac_ctx* c) {
    const double rl = c->bc.rc + c->skin;
    for (int d = 0; d < 3; ++d) {
        int n = (int)floor(kHalf * c->bc.box[d] / (rl * (1.0 + 1e-9)));
        c->ncell[d] = n < 1 ? 1 : n;
    }
    c->n_cell = c->ncell[0] * c->ncell[1] * c->ncell[2];
    Stencil st;
    make_stencil(c->ncell, &st);
    if (!c->stencil) {
        if (hipMalloc(&c->stencil, sizeof(Stencil)) != hipSuccess) {
            capac_set_error("stencil allocation failed");
            return CAPAC_ERR_HIP;
        }
    }
    if (hipMemcpy(c->stencil, &st, sizeof(Stencil), hipMemcpyHostToDevice) != hipSuccess) {
        capac_set_error("stencil upload failed");
        return CAPAC_ERR_HIP;
    }
    return 0;
}

static int nb_launch(capac_ctx* c) {
    const int N = c->bc.N;
    hipStream_t s = c->stream;
    const unsigned long long* need = c->nb_ctl;
    int tb = 256, gb = (N + tb - 1) / tb;
    k_cell_assign<<<gb, tb, 0, s>>>(c->r, c->bc, c->ncell[0], c->ncell[1], c->ncell[2], c->n_cell,
                                    c->cell_of, c->cell_rank, c->cell_count, need);
    k_cell_scan<<<1, 1024, 0, s>>>(c->cell_count, c->n_cell, c->cell_start, c->cell_fill, need);
    k_cell_fill<<<gb, tb, 0, s>>>(c->cell_of, c->cell_rank, N, c->n_cell, c->cell_fill, c->cell_tmp, need);
    const double3 box = make_double3(c->bc.box[0], c->bc.box[1], c->bc.box[2]);
    k_cell_sort<<<c->n_cell, 256, 0, s>>>(c->cell_start, c->cell_tmp, c->r, c->mol_id, c->local_idx, N,
                                          c->cell_atoms, c->sorted_r, c->sorted_r + N,
                                          c->sorted_r + 2 * N, c->sorted_mol, c->sorted_loc, box, need);
    bool shifted = true;
    for (int d = 0; d < 3; ++d) shifted = shifted && c->ncell[d] >= 2 * kHalf + 1;
    auto build = shifted ? k_nb_build<true> : k_nb_build<false>;
    build<<<c->n_cell, kBuildThreads, 0, s>>>(c->sorted_r, c->sorted_r + N, c->sorted_r + 2 * N,
                                            c->sorted_mol, c->sorted_loc, c->cell_atoms, c->cell_start,
                                            c->excl, c->bc, c->ncell[0], c->ncell[1], c->ncell[2],
                                            (const Stencil*)c->stencil, c->nb_stride, c->n_front,
                                            c->n_back, c->nbr, c->nb_overflow, c->tail_list, need);
    k_mark_built<<<(3 * N + 255) / 256, 256, 0, s>>>(c->r, c->r_build, 3 * N, c->nb_ctl);
    CHECK_HIP(hipGetLastError());
    return 0;
}

// unconditional build
int nb_build(capac_ctx* c) {
    k_set_flag<<<1, 64, 0, c->stream>>>(c->nb_ctl);
    k_decide<<<1, 256, 0, c->stream>>>(c->nb_ctl, 0.0, c->cell_count, c->n_cell, c->nb_overflow, c->tail_list);
    return nb_launch(c);
}

// build only when an atom moved more than skin/2 since the last build (always
// without a skin)
int nb_build_if_needed(capac_ctx* c) {
    if (c->skin <= 0.0) return nb_build(c);
    const int N = c->bc.N;
    // half the skin, less a relative 1e-6 so rounding in the distance tests
    // cannot matter at the boundary
    double half = 0.5 * c->skin * (1.0 - 1e-6);
    k_disp_decide<<<(N + kDispThreads - 1) / kDispThreads, kDispThreads, 0, c->stream>>>(c->r, c->r_build, N, c->nb_ctl, half * half,
                                                          c->cell_count, c->n_cell, c->nb_overflow,
                                                          c->tail_list);
    return nb_launch(c);
}
