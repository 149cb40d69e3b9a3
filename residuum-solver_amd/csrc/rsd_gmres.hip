// rsd_gmres.hip -- right-preconditioned restarted GMRES(m) on the device
// (linalg/solver/iterative/gmres/Solver.tpp:5-170).
//
// Modified Gram-Schmidt runs as one fused kernel per basis vector: it
// finishes the previous projection w -= h_{i-1,j} q_{i-1} element by element
// and accumulates q_i . w for the next coefficient, which is exactly the
// reference's dot-then-axpy sequence.  The Hessenberg column, the Givens
// rotations and the back-substitution are scalar work done by one device
// thread.  A whole cycle is launched at once: after the stop flag is set
// the remaining kernels of the cycle return immediately, and the host reads
// the state once per cycle.
#include "rsd_gmres.hpp"
#include "rsd_linalg.hpp"
#include "rsd_rowdot.hpp"

#include <algorithm>
#include <cmath>

namespace rsd {

// w = A (M q): the Jacobi product q_c * invd_c is recomputed at every
// gathered column, bitwise the vector z = M q the reference forms first
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_gm_apply(int64_t n, int64_t nb, Cols C, const double* __restrict__ val,
                                                     const double* __restrict__ q, const double* __restrict__ invd,
                                                     double* __restrict__ w, const GMState* st) {
    if (st->stop) return;
    const int64_t r = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t base = C.slice_ptr[r >> 6] + (r & 63);
    const RowCols<MODE> rc(r, C);
    double sum = 0.0;
    for (int j = 0; j < rc.len; ++j) {
        const int64_t s = base + (int64_t)j * kSlice;
        const int c = rc.col(C, r, s, j);
        const double z = invd ? q[c] * invd[c] : q[c];
        sum += stream_load(val + s) * z;
    }
    w[r] = sum;
}

// MGS step i of Arnoldi column j: block partials of q_i . w (of w . w for
// the norm step, Qi == nullptr)
__global__ __launch_bounds__(kBlock) void k_gm_mgs(int64_t n, int64_t nb, double* __restrict__ w,
                                                   const double* __restrict__ Qprev, const double* __restrict__ Qi,
                                                   int32_t i, int32_t j, int32_t m, const double* H, const GMState* st,
                                                   double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t k = lb * kBlock + threadIdx.x;
    const double hprev = i > 0 ? H[(int64_t)j * (m + 1) + (i - 1)] : 0.0;
    double v[1] = {0.0};
    if (k < n) {
        double wk = w[k];
        if (i > 0) {
            wk = wk + -hprev * Qprev[k];
            w[k] = wk;
        }
        v[0] = Qi ? Qi[k] * wk : wk * wk;
    }
    block_partials<1>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kFold) void k_gm_mgs_fold(int64_t nb, const double* __restrict__ partials, int32_t i,
                                                        int32_t j, int32_t m, int32_t norm, double* H, GMState* st) {
    if (st->stop) return;
    double out[1];
    fold_partials<1>(nb, partials, out);
    if (threadIdx.x != 0) return;
    if (norm) st->hnext = sqrt(out[0]);
    else H[(int64_t)j * (m + 1) + i] = out[0];
}

// q_{j+1} = w * (1 / hnext) unless the column broke down
__global__ void k_gm_newq(int64_t n, const double* __restrict__ w, double* __restrict__ q, const GMState* st) {
    if (st->stop || st->hnext <= 1e-14) return;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        const double s = 1.0 / st->hnext;
        q[k] = w[k] * s;
    }
}

// Givens rotations for column j (Solver.tpp:78-122)
__global__ void k_gm_givens(int32_t j, int32_t m, double* H, double* cs, double* sn, double* g, GMState* st,
                            double tol, int32_t abs_tol, int64_t max_iters) {
    if (st->stop) return;
    auto hat = [&](int i, int jj) -> double& { return H[(int64_t)jj * (m + 1) + i]; };
    const double hnext = st->hnext;
    st->total += 1;
    const int32_t breakdown = hnext <= 1e-14;
    for (int i = 0; i < j; ++i) {
        const double h1 = hat(i, j), h2 = hat(i + 1, j);
        hat(i, j) = cs[i] * h1 + sn[i] * h2;
        hat(i + 1, j) = -sn[i] * h1 + cs[i] * h2;
    }
    const double hjj = hat(j, j);
    const double den = sqrt(hjj * hjj + hnext * hnext);
    if (den <= 1e-14) { cs[j] = 1.0; sn[j] = 0.0; }
    else { cs[j] = hjj / den; sn[j] = hnext / den; }
    hat(j, j) = cs[j] * hjj + sn[j] * hnext;
    const double gj = g[j];
    g[j] = cs[j] * gj;
    g[j + 1] = -sn[j] * gj;
    const double est = fabs(g[j + 1]);
    st->est = est;
    st->mloc = j + 1;
    st->breakdown = breakdown;
    if ((abs_tol ? est : est / (st->res0 + 1e-50)) < tol || breakdown || st->total >= max_iters) st->stop = 1;
}

// back-substitution H y = g (Solver.tpp:126-133)
__global__ void k_gm_backsub(int32_t m, const double* H, const double* g, double* y, const GMState* st) {
    const int mloc = st->mloc;
    for (int ii = 0; ii < mloc; ++ii) {
        const int i = mloc - 1 - ii;
        double s = g[i];
        for (int k = i + 1; k < mloc; ++k) s -= H[(int64_t)k * (m + 1) + i] * y[k];
        y[i] = s / H[(int64_t)i * (m + 1) + i];
    }
}

// x += M (sum_i y_i q_i)  (Solver.tpp:135-139)
__global__ void k_gm_update_x(int64_t n, const double* const* __restrict__ Q, const double* __restrict__ y,
                              const double* __restrict__ invd, double* __restrict__ x, const GMState* st) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    double z = 0.0;
    for (int i = 0; i < st->mloc; ++i) z += y[i] * Q[i][k];
    const double wk = invd ? z * invd[k] : z;
    x[k] += 1.0 * wk;
}

// q0 = b - A x, beta = ||q0|| (Solver.tpp:22-27, 141-147)
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_gm_residual(int64_t n, int64_t nb, Cols C, const double* __restrict__ val,
                                                        const double* __restrict__ b, const double* __restrict__ x,
                                                        double* __restrict__ q0, double* partials) {
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t k = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (k < n) {
        const double qk = b[k] + -1.0 * row_dot_f64<MODE>(k, C, val, x);
        q0[k] = qk;
        v[0] = qk * qk;
    }
    block_partials<1>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kFold) void k_gm_residual_fold(int64_t nb, const double* __restrict__ partials,
                                                             GMState* st, int32_t first) {
    double out[1];
    fold_partials<1>(nb, partials, out);
    if (threadIdx.x == 0) {
        const double beta = sqrt(out[0]);
        st->beta = beta;
        if (first) st->res0 = beta;
    }
}

// start of a cycle: q0 *= 1/beta, g = (beta, 0, ...), clear the stop flag
__global__ void k_gm_cycle_start(int64_t n, double* __restrict__ q0, double* g, int32_t m, GMState* st) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double beta = st->beta;
    if (k < n) q0[k] *= 1.0 / beta;
    if (k == 0) {
        g[0] = beta;
        for (int i = 1; i <= m; ++i) g[i] = 0.0;
    }
}

__global__ void k_gm_reset_flags(GMState* st) {
    st->stop = 0;
    st->breakdown = 0;
    st->mloc = 0;
}

GMRESResult gmres_run(rsd_ctx* ctx, const rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                      int32_t m, int32_t precond, bool abs_tol) {
    auto crit = [&](double res, double res0) { return abs_tol ? res : res / (res0 + 1e-50); };
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = std::max<int64_t>(1, nblocks(n)), nbe = nblocks(std::max<int64_t>(n, 1));
    hipStream_t st = ctx->stream;
    const size_t vb = sizeof(double) * (size_t)std::max<int64_t>(n, 1);
    std::vector<double*> hq(m + 1);
    for (int i = 0; i <= m; ++i) hq[i] = (double*)ctx->ws("gm.Q" + std::to_string(i), vb);
    double* w = (double*)ctx->ws("gm.w", vb);
    double* invd = precond == RSD_PC_JACOBI ? (double*)ctx->ws("gm.invd", vb) : nullptr;
    double* H = (double*)ctx->ws("gm.H", sizeof(double) * (size_t)(m + 1) * m);
    double* cs = (double*)ctx->ws("gm.cs", sizeof(double) * (size_t)(m + 1));
    double* sn = (double*)ctx->ws("gm.sn", sizeof(double) * (size_t)(m + 1));
    double* g = (double*)ctx->ws("gm.g", sizeof(double) * (size_t)(m + 1));
    double* y = (double*)ctx->ws("gm.y", sizeof(double) * (size_t)(m + 1));
    double** dQ = (double**)ctx->ws("gm.Qptr", sizeof(double*) * (size_t)(m + 1));
    GMState* dst = (GMState*)ctx->ws("gm.state", sizeof(GMState));
    double* part = ctx->partials(nb);
    RSD_HIP(hipMemcpyAsync(dQ, hq.data(), sizeof(double*) * (size_t)(m + 1), hipMemcpyHostToDevice, st));
    RSD_HIP(hipMemsetAsync(dst, 0, sizeof(GMState), st));
    if (invd) launch_inv_diag(ctx, A, invd);
    GMState* h = nullptr;
    RSD_HIP(hipHostMalloc((void**)&h, sizeof(GMState), hipHostMallocDefault));
    auto read = [&] {
        RSD_HIP(hipMemcpyAsync(h, dst, sizeof(GMState), hipMemcpyDeviceToHost, st));
        RSD_HIP(hipStreamSynchronize(st));
    };
    GMRESResult res{};
    try {
        RSD_LAUNCH_COLS(S.mode(), k_gm_residual, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(), A->val.p, b, x,
                        hq[0], part);
        hipLaunchKernelGGL(k_gm_residual_fold, dim3(1), dim3(kFold), 0, st, nb, part, dst, 1);
        read();
        const double res0 = h->res0;
        res.initial_residual = res0;
        if (crit(res0, res0) < tol) {
            res.converged = 1;
            res.final_residual = res0;
            res.final_residual_rel = 1.0;
            res.iterations = 0;
        } else {
            int64_t total = 0;
            for (;;) {
                hipLaunchKernelGGL(k_gm_reset_flags, dim3(1), dim3(1), 0, st, dst);
                hipLaunchKernelGGL(k_gm_cycle_start, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, hq[0], g, m, dst);
                for (int j = 0; j < m; ++j) {
                    RSD_LAUNCH_COLS(S.mode(), k_gm_apply, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(),
                                    A->val.p, hq[j], invd, w, dst);
                    for (int i = 0; i <= j + 1; ++i) {
                        const bool norm = i == j + 1;
                        hipLaunchKernelGGL(k_gm_mgs, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb, w,
                                           i > 0 ? hq[i - 1] : nullptr, norm ? nullptr : hq[i], i, j, m, H, dst, part);
                        hipLaunchKernelGGL(k_gm_mgs_fold, dim3(1), dim3(kFold), 0, st, nb, part, i, j, m,
                                           (int32_t)norm, H, dst);
                    }
                    hipLaunchKernelGGL(k_gm_newq, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, w, hq[j + 1], dst);
                    hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(1), 0, st, j, m, H, cs, sn, g, dst, tol,
                                       (int32_t)abs_tol, max_iters);
                    RSD_HIP(hipGetLastError());
                }
                hipLaunchKernelGGL(k_gm_backsub, dim3(1), dim3(1), 0, st, m, H, g, y, dst);
                hipLaunchKernelGGL(k_gm_update_x, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, (const double* const*)dQ,
                                   y, invd, x, dst);
                RSD_LAUNCH_COLS(S.mode(), k_gm_residual, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(),
                                A->val.p, b, x, hq[0], part);
                hipLaunchKernelGGL(k_gm_residual_fold, dim3(1), dim3(kFold), 0, st, nb, part, dst, 0);
                RSD_HIP(hipGetLastError());
                read();
                total = h->total;
                const double beta = h->beta;
                const bool conv = crit(beta, res0) < tol;
                if (conv || h->breakdown || total >= max_iters) {
                    res.converged = conv;
                    res.final_residual = beta;
                    res.final_residual_rel = beta / (res0 + 1e-50);
                    res.iterations = total;
                    break;
                }
            }
        }
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
    return res;
}

}  // namespace rsd

using namespace rsd;

extern "C" int rsd_gmres_solve_f64(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol,
                                   int32_t tol_type, int64_t max_iters, int32_t krylov, int32_t precond,
                                   rsd_report* rep) {
    return guard([&] {
        RSD_CHECK(ctx && A && b && x, "rsd_gmres_solve_f64: null argument");
        RSD_CHECK(krylov >= 1 && krylov <= 512, "krylov dimension outside [1, 512]");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        RSD_CHECK(tol_type == RSD_TOL_RELATIVE || tol_type == RSD_TOL_ABSOLUTE, "unknown tolerance type");
        GMRESResult r = gmres_run(ctx, A, b, x, tol, max_iters, krylov, precond, tol_type == RSD_TOL_ABSOLUTE);
        if (rep) {
            rep->converged = r.converged;
            rep->iterations = r.iterations;
            rep->initial_residual = r.initial_residual;
            rep->final_residual = r.final_residual;
            rep->final_residual_rel = r.final_residual_rel;
        }
    });
}
