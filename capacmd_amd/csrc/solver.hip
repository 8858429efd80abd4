// Jacobi-preconditioned BiCGSTAB on the relay matrix: the reference's
// mpi_precon_bicg_stab_COO (bicg_stab.c:253-573) with its exact control flow:
//   bTb < 1e-10 -> return, x unchanged                       (bicg_stab.c:300)
//   criteria = 1e-8 bTb, warm start from the resident x       (:303-306)
//   rho < 1e-10 at the first iteration -> return             (:368)
//   omega = (Kt, Ks)/(Kt, Kt)                                 (:483-488)
//   stop when rTr < criteria; after 1000 iterations accept rTr < 100 criteria
//   ("loosened"), else fail                                   (:539-571)
// Lagrange rows/columns of the neutrality constraint: the charge row of metal
// i gets + lambda_k(i); the row of nanoparticle k is sum_{i in k} q_i; their
// preconditioner entries are 1.
//
// Per iteration: two fused SpMV+dot launches (one workgroup per metal
// row-block, streaming its 16 x Nm block components once) and two
// single-workgroup vector kernels that finish the dots in a fixed order and
// update the vectors (the second also opens the next iteration).  All scalars
// live on the device; a `done` flag turns the remaining launches of a batch
// into no-ops, so the host synchronises once per batch, not per iteration.
// Cross-workgroup hand-off stays at kernel boundaries: on this GPU the eight
// XCDs have separate L2 caches, and an in-kernel device-scope fence costs
// more than a launch.
#include "capac_internal.h"

#include <stdlib.h>

struct Solver {
    int n_mat, Nm, n_np;
    double *r0, *r, *p, *v, *s, *t, *y, *z, *K, *Ax;
    double* part;           // (Nm + n_np) * 2 partial dots
    SolveScalars* sc;       // device
    SolveScalars* sc_host;  // pinned
};

namespace {

constexpr int kVecThreads = 1024;

__device__ double block_sum(double v, double* s_w) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) s_w[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += s_w[w];
    return t;
}


// y = A x for the 4 Nm relay rows, plus partial dots; blocks >= Nm compute
// the Lagrange rows.  mode 0: no dots; 1: part[i] = (r0, y) rows of i;
// 2: part[2i] = (K y, x), part[2i+1] = (K y, K y)
__global__ void __launch_bounds__(256)
k_spmv64(const double* __restrict__ relay, int ld, int Nm, int n_np, const int* __restrict__ np_of,
         const int* __restrict__ np_first, const int* __restrict__ np_last,
         const double* __restrict__ x, double* __restrict__ y, int mode,
         const double* __restrict__ r0, const double* __restrict__ K, double* __restrict__ part,
         const SolveScalars* __restrict__ sc) {
    __shared__ double s_w[4][4];
    if (sc && sc->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int i = blockIdx.x;
    if (i >= Nm) {
        // Lagrange row of nanoparticle k: sum of its metal charges
        int k = i - Nm;
        double acc = 0.0;
        for (int m = np_first[k] + threadIdx.x; m <= np_last[k]; m += blockDim.x) acc += x[3 * Nm + m];
        acc = wave_sum(acc);
        if (lane == 0) s_w[wave][0] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = s_w[0][0] + s_w[1][0] + s_w[2][0] + s_w[3][0];
            int row = 4 * Nm + k;
            y[row] = t;
            if (mode == 1) part[i] = r0[row] * t;
            if (mode == 2) { part[2 * i] = t * x[row]; part[2 * i + 1] = t * t; }
        }
        return;
    }
    const double* base = relay + (size_t)i * 16 * ld;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int j = threadIdx.x; j < Nm; j += blockDim.x) {
        double xj0 = x[3 * j], xj1 = x[3 * j + 1], xj2 = x[3 * j + 2], xj3 = x[3 * Nm + j];
        const double* bp = base + j;
        a0 += bp[0 * ld] * xj0 + bp[1 * ld] * xj1 + bp[2 * ld] * xj2 + bp[3 * ld] * xj3;
        a1 += bp[4 * ld] * xj0 + bp[5 * ld] * xj1 + bp[6 * ld] * xj2 + bp[7 * ld] * xj3;
        a2 += bp[8 * ld] * xj0 + bp[9 * ld] * xj1 + bp[10 * ld] * xj2 + bp[11 * ld] * xj3;
        a3 += bp[12 * ld] * xj0 + bp[13 * ld] * xj1 + bp[14 * ld] * xj2 + bp[15 * ld] * xj3;
    }
    a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2); a3 = wave_sum(a3);
    if (lane == 0) { s_w[wave][0] = a0; s_w[wave][1] = a1; s_w[wave][2] = a2; s_w[wave][3] = a3; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[4];
        for (int c = 0; c < 4; ++c) o[c] = s_w[0][c] + s_w[1][c] + s_w[2][c] + s_w[3][c];
        int np = np_of[i];
        if (np >= 0) o[3] += x[4 * Nm + np];
        int rows[4] = {3 * i, 3 * i + 1, 3 * i + 2, 3 * Nm + i};
        for (int c = 0; c < 4; ++c) y[rows[c]] = o[c];
        if (mode == 1) {
            double d = 0.0;
            for (int c = 0; c < 4; ++c) d += r0[rows[c]] * o[c];
            part[i] = d;
        } else if (mode == 2) {
            double d0 = 0.0, d1 = 0.0;
            for (int c = 0; c < 4; ++c) {
                double kt = K[rows[c]] * o[c];
                d0 += kt * x[rows[c]];
                d1 += kt * kt;
            }
            part[2 * i] = d0;
            part[2 * i + 1] = d1;
        }
    }
}

// start of BiCGSTAB (bicg_stab.c:294-404): bTb, criteria, r0 = r = b - Ax,
// v = 0, K = 1/diag, rho = (r0, r) with the two early exits, and the first
// direction p = r + rho (0 - 0) = r, y = K p
__global__ void __launch_bounds__(kVecThreads)
k_bicg_init(int n_mat, int n4, const double* __restrict__ b, const double* __restrict__ Ax,
            const double* __restrict__ diag, double* __restrict__ r0, double* __restrict__ r,
            double* __restrict__ p, double* __restrict__ v, double* __restrict__ K,
            double* __restrict__ y, SolveScalars* sc) {
    __shared__ double s_w[16];
    double bb = 0.0, rr = 0.0;
    for (int k = threadIdx.x; k < n_mat; k += blockDim.x) {
        double bk = b[k];
        double rk = bk - Ax[k];
        r0[k] = rk;
        r[k] = rk;
        v[k] = 0.0;
        K[k] = k < n4 ? 1.0 / diag[k] : 1.0;
        bb += bk * bk;
        rr += rk * rk;
    }
    double bTb = block_sum(bb, s_w);
    double rho = block_sum(rr, s_w);
    const bool zero_rhs = bTb < 1.0e-10;
    const bool zero_rho = !zero_rhs && rho < 1.0e-10;
    if (!zero_rhs && !zero_rho)
        for (int k = threadIdx.x; k < n_mat; k += blockDim.x) {
            double pk = r[k];
            p[k] = pk;
            y[k] = pk * K[k];
        }
    if (threadIdx.x == 0) {
        sc->bTb = bTb;
        sc->criteria = bTb * 1.0e-08;
        sc->rho_old = 1.0;
        sc->alpha = 1.0;
        sc->omega = 1.0;
        sc->rho = rho;
        sc->rTr = 0.0;
        sc->iter = 0;
        sc->done = (zero_rhs || zero_rho) ? 1 : 0;
        sc->kind = zero_rhs ? CAPAC_SOLVE_ZERO_RHS : (zero_rho ? CAPAC_SOLVE_ZERO_RHO : CAPAC_SOLVE_CONVERGED);
    }
}

// alpha = rho / (r0, v); s = r - alpha v; z = K s
__global__ void __launch_bounds__(kVecThreads)
k_bicg_s(int n_mat, int n_part, const double* __restrict__ part, const double* __restrict__ r,
         const double* __restrict__ v, const double* __restrict__ K, double* __restrict__ s,
         double* __restrict__ z, SolveScalars* sc) {
    __shared__ double s_w[16];
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = sc->done;
    __syncthreads();
    if (s_skip) return;
    double acc = 0.0;
    for (int k = threadIdx.x; k < n_part; k += blockDim.x) acc += part[k];
    double r0v = block_sum(acc, s_w);
    double alpha = sc->rho / r0v;
    for (int k = threadIdx.x; k < n_mat; k += blockDim.x) {
        double sk = r[k] - alpha * v[k];
        s[k] = sk;
        z[k] = sk * K[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) sc->alpha = alpha;
}

// omega = (Kt, z) / (Kt, Kt); x += alpha y + omega z; r = s - omega t; rTr and
// the stop tests (bicg_stab.c:483-571); when not done, the next iteration's
// rho = (r0, r), beta = (rho/rho_old)(alpha/omega), p = r + beta (p - omega v)
// and y = K p in the same pass (bicg_stab.c:365-404)
__global__ void __launch_bounds__(kVecThreads)
k_bicg_xp(int n_mat, int n_part, const double* __restrict__ part, double* __restrict__ x,
          double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ s,
          const double* __restrict__ t, double* __restrict__ r, const double* __restrict__ r0,
          double* __restrict__ p, const double* __restrict__ v, const double* __restrict__ K,
          SolveScalars* sc) {
    __shared__ double s_w[16];
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = sc->done;
    __syncthreads();
    if (s_skip) return;
    double a0 = 0.0, a1 = 0.0;
    for (int k = threadIdx.x; k < n_part; k += blockDim.x) { a0 += part[2 * k]; a1 += part[2 * k + 1]; }
    double ktks = block_sum(a0, s_w);
    double ktkt = block_sum(a1, s_w);
    double omega = ktks / ktkt;
    double alpha = sc->alpha;
    double rho = sc->rho;
    double acc = 0.0, acc_r0 = 0.0;
    for (int k = threadIdx.x; k < n_mat; k += blockDim.x) {
        x[k] += alpha * y[k] + omega * z[k];
        double rk = s[k] - omega * t[k];
        r[k] = rk;
        acc += rk * rk;
        acc_r0 += r0[k] * rk;
    }
    double rTr = block_sum(acc, s_w);
    double rho_next = block_sum(acc_r0, s_w);
    const int iter = sc->iter + 1;
    int done = 0, kind = CAPAC_SOLVE_CONVERGED;
    if (rTr < sc->criteria) {
        done = 1;
    } else if (iter > 1000) {
        done = 1;
        kind = rTr < sc->criteria * 100.0 ? CAPAC_SOLVE_LOOSENED : CAPAC_SOLVE_FAILED;
    }
    if (!done) {
        double beta = (rho_next / rho) * (alpha / omega);
        for (int k = threadIdx.x; k < n_mat; k += blockDim.x) {
            double pk = r[k] + beta * (p[k] - omega * v[k]);
            p[k] = pk;
            y[k] = pk * K[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sc->omega = omega;
        sc->rho_old = rho;
        sc->rho = rho_next;
        sc->iter = iter;
        sc->rTr = rTr;
        sc->done = done;
        sc->kind = kind;
    }
}

}  // namespace

int solver_create(capac_ctx* c) {
    Solver* S = (Solver*)calloc(1, sizeof(Solver));
    if (!S) { capac_set_error("out of host memory"); return CAPAC_ERR_ARG; }
    c->solver = S;
    S->n_mat = c->n_mat;
    S->Nm = c->bc.Nm;
    S->n_np = c->bc.n_np;
    size_t n = (size_t)(c->n_mat > 0 ? c->n_mat : 1);
    double** vecs[10] = {&S->r0, &S->r, &S->p, &S->v, &S->s, &S->t, &S->y, &S->z, &S->K, &S->Ax};
    for (int k = 0; k < 10; ++k) {
        CHECK_HIP(hipMalloc(vecs[k], n * sizeof(double)));
        CHECK_HIP(hipMemset(*vecs[k], 0, n * sizeof(double)));
    }
    CHECK_HIP(hipMalloc(&S->part, 2 * (size_t)(S->Nm + S->n_np + 1) * sizeof(double)));
    CHECK_HIP(hipMalloc(&S->sc, sizeof(SolveScalars)));
    CHECK_HIP(hipMemset(S->sc, 0, sizeof(SolveScalars)));
    CHECK_HIP(hipHostMalloc(&S->sc_host, sizeof(SolveScalars), hipHostMallocDefault));
    return 0;
}

void solver_destroy(capac_ctx* c) {
    Solver* S = c->solver;
    if (!S) return;
    double* vecs[10] = {S->r0, S->r, S->p, S->v, S->s, S->t, S->y, S->z, S->K, S->Ax};
    for (int k = 0; k < 10; ++k) (void)hipFree(vecs[k]);
    (void)hipFree(S->part);
    (void)hipFree(S->sc);
    (void)hipHostFree(S->sc_host);
    free(S);
    c->solver = NULL;
}

static void spmv(capac_ctx* c, const double* x, double* y, int mode, const SolveScalars* sc) {
    Solver* S = c->solver;
    k_spmv64<<<S->Nm + S->n_np, 256, 0, c->stream>>>(c->relay, c->relay_ld, S->Nm, S->n_np,
                                                     c->np_of_metal, c->np_first, c->np_last, x, y,
                                                     mode, S->r0, S->K, S->part, sc);
}

// Solves relay * x = b with the resident matrix; x is the warm start and the
// result.  Fills c->solve_*; returns CAPAC_ERR_SOLVER when the reference
// would exit (bicg_stab.c:563-569).
int solver_run(capac_ctx* c, const double* b, double* x) {
    Solver* S = c->solver;
    const int n = S->n_mat, n4 = 4 * S->Nm, n_part = S->Nm + S->n_np;
    hipStream_t st = c->stream;
    spmv(c, x, S->Ax, 0, NULL);
    k_bicg_init<<<1, kVecThreads, 0, st>>>(n, n4, b, S->Ax, c->diag, S->r0, S->r, S->p, S->v, S->K,
                                           S->y, S->sc);
    // first batch: the previous solve's iteration count (warm starts change
    // slowly), then small increments, so few launches are wasted as no-ops
    const int first = c->solve_iters > 2 ? c->solve_iters : 4;
    int batch = first, launched = 0;
    for (;;) {
        for (int it = 0; it < batch; ++it) {
            spmv(c, S->y, S->v, 1, S->sc);
            k_bicg_s<<<1, kVecThreads, 0, st>>>(n, n_part, S->part, S->r, S->v, S->K, S->s, S->z,
                                                S->sc);
            spmv(c, S->z, S->t, 2, S->sc);
            k_bicg_xp<<<1, kVecThreads, 0, st>>>(n, n_part, S->part, x, S->y, S->z, S->s, S->t, S->r,
                                                 S->r0, S->p, S->v, S->K, S->sc);
        }
        launched += batch;
        CHECK_HIP(hipGetLastError());
        CHECK_HIP(hipMemcpyAsync(S->sc_host, S->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, st));
        CHECK_HIP(hipStreamSynchronize(st));
        if (S->sc_host->done) break;
        if (launched > 1010) {
            capac_set_error("BiCGSTAB did not report completion after %d iterations", launched);
            return CAPAC_ERR_SOLVER;
        }
        batch = (launched == first) ? 2 : (batch < 16 ? 2 * batch : 16);
    }
    c->solve_iters = S->sc_host->iter;
    c->solve_kind = S->sc_host->kind;
    c->solve_rTr = S->sc_host->rTr;
    c->solve_bTb = S->sc_host->bTb;
    c->timing.solve_iters += S->sc_host->iter;
    if (c->solve_kind == CAPAC_SOLVE_FAILED) {
        capac_set_error("Error: bicg_stab not converged. rTr= %e, bTb= %e", c->solve_rTr, c->solve_bTb);
        return CAPAC_ERR_SOLVER;
    }
    return 0;
}

// one plain SpMV over the resident matrix (for kernel timing)
int solver_spmv_only(capac_ctx* c, const double* x, double* y) {
    spmv(c, x, y, 0, NULL);
    CHECK_HIP(hipGetLastError());
    return 0;
}
