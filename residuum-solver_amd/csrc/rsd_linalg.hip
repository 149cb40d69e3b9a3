// rsd_linalg.hip -- fills, BLAS-1, sliced-ELL SpMV, Jacobi and fused PCG.
//
// Products and sums are rounded separately (built with -ffp-contract=off):
// the SpMV row sums reproduce the reference's left-to-right order
// (linalg/operations/backend/cpu/MatrixOps.tpp:5-23) bit for bit, and every
// reduction uses the fixed block tree of rsd_internal.hpp.
#include "rsd_linalg.hpp"
#include "rsd_rowdot.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace rsd {

// ------------------------------------------------------------------------
// elementwise kernels
// ------------------------------------------------------------------------
template<typename T>
__global__ void k_fill(int64_t n, T* __restrict__ p, T v) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

__global__ void k_axpby(int64_t n, double a, double b, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = b * y[i] + a * x[i];
}

__global__ void k_scal(int64_t n, double a, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] *= a;
}

__global__ void k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = x[i];
}

__global__ void k_mul(int64_t n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) z[i] = x[i] * y[i];
}

// FP32 vectors (Vector<float>): the elementwise operations in float, each
// product and sum rounded once as in the CPU loop
__global__ void k_axpy_f32(int64_t n, float a, const float* __restrict__ x, float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

__global__ void k_axpby_f32(int64_t n, float a, float b, const float* __restrict__ x, float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = b * y[i] + a * x[i];
}

__global__ void k_scal_f32(int64_t n, float a, float* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] *= a;
}

__global__ void k_copy_f32(int64_t n, const float* __restrict__ x, float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = x[i];
}

__global__ void k_mul_f32(int64_t n, const float* __restrict__ x, const float* __restrict__ y, float* __restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) z[i] = x[i] * y[i];
}

// FP32 dot: each product exact in FP64, summed with the FP64 tree
__global__ __launch_bounds__(kBlock) void k_dot_f32(int64_t n, int64_t nb, const float* __restrict__ x,
                                                    const float* __restrict__ y, double* partials) {
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t i = lb * kBlock + threadIdx.x;
    double v[1] = {i < n ? (double)x[i] * (double)y[i] : 0.0};
    block_partials<1>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, int64_t nb, const double* __restrict__ x,
                                                const double* __restrict__ y, double* partials) {
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t i = lb * kBlock + threadIdx.x;
    double v[1] = {i < n ? x[i] * y[i] : 0.0};
    block_partials<1>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kFold) void k_dot_fold(int64_t nb, const double* __restrict__ partials, double* out,
                                                    int32_t take_sqrt) {
    double r[1];
    fold_partials<1>(nb, partials, r);
    if (threadIdx.x == 0) *out = take_sqrt ? sqrt(r[0]) : r[0];
}

// ------------------------------------------------------------------------
// SpMV: one thread per row, one wavefront per 64-row slice
// ------------------------------------------------------------------------
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_f64(int64_t n, int64_t nb, Cols C, const double* __restrict__ val,
                                                     const double* __restrict__ x, double* __restrict__ y) {
    const int64_t r = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    if (r < n) y[r] = row_dot_f64<MODE>(r, C, val, x);
}

// The stencil copy from the sliced values: row r (one thread) writes all
// its NP positions, its entries in CSR order at the positions its mask
// holds, zeros elsewhere, as 16-byte pairs.  The copy's slots of the rows
// past n stay as reserved (zero).
template<int NP>
__global__ __launch_bounds__(kBlock) void k_to_st(int64_t n, const int64_t* __restrict__ slice_ptr,
                                                  const uint32_t* __restrict__ mask, const double* __restrict__ val,
                                                  double* __restrict__ vst) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const uint32_t m = mask[r];
    const int last = __builtin_popcount(m) - 1;
    const double* src = val + slice_ptr[r >> 6] + (r & 63);
    double v[NP];
    int j = 0;
#pragma unroll
    for (int b = 0; b < NP; ++b) {
        const bool has = (m >> b) & 1u;
        // entry min(j, last) is one of the row's own (an empty row loads none)
        const double t = last >= 0 ? src[(int64_t)min(j, last) * kSlice] : 0.0;
        v[b] = has ? t : 0.0;
        j += has;
    }
    double* dst = vst + (r >> 6) * (NP * 64);
    const int64_t l = r & 63;
#pragma unroll
    for (int b = 0; b + 1 < NP; b += 2) *(rsd_double2*)(dst + 128 * (b >> 1) + 2 * l) = rsd_double2{v[b], v[b + 1]};
    dst[128 * ((NP - 1) >> 1) + l] = v[NP - 1];
}

// y = A x over the stencil copy.  Every lane runs the row product (the
// lanes past n feed their neighbours, row_dot_st); only rows < n store.
template<int NR>
__global__ __launch_bounds__(kBlock) void k_spmv_st(int64_t n, int64_t nb, StCols S, const double* __restrict__ val,
                                                    const double* __restrict__ x, double* __restrict__ y) {
    const int64_t r = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    const double s = row_dot_st<NR>(r, n, S, val, x);
    if (r < n) y[r] = s;
}

__global__ void k_inv_diag(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ val,
                           double* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t s = diag[i];
    const double d = s >= 0 ? val[s] : 0.0;
    inv[i] = (fabs(d) > 1e-300) ? 1.0 / d : 1.0;
}

// ------------------------------------------------------------------------
// fused PCG (cg/Solver.tpp:15-119), five launches per iteration:
//   spmv_pap:  Ap = A p, block sums of p.Ap
//   alpha:     (one block) alpha = rz_old / pAp
//   update_r:  x += alpha p, r += (-alpha) Ap, z = M r, block sums of r.r, r.z
//   beta:      (one block) ||r||, convergence, beta, rz_old
//   update_p:  p = beta p + z
// The state lives on the device; once `stop` is set every kernel returns at
// once, so fixed-size chunks (or graph replays) reproduce the reference's
// stopping iterate exactly.  x is updated in update_r, before the
// convergence test of the same iteration, as cg/Solver.tpp:40-44 does.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cg_init(int64_t n, int64_t nb, const double* __restrict__ b,
                                                    const double* __restrict__ Ax, double* __restrict__ r,
                                                    double* __restrict__ p, const double* __restrict__ invd,
                                                    double* partials) {
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t i = lb * kBlock + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        const double ri = b[i] + -1.0 * Ax[i];
        const double zi = invd ? ri * invd[i] : ri;
        r[i] = ri;
        p[i] = zi;
        v[0] = ri * ri;
        v[1] = ri * zi;
    }
    block_partials<2>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kFold) void k_cg_init_fold(int64_t nb, const double* __restrict__ partials, CGState* st,
                                                        double tol, int32_t abs_tol, int64_t max_iters, double* hist) {
    double out[2];
    fold_partials<2>(nb, partials, out);
    if (threadIdx.x != 0) return;
    const double res0 = sqrt(out[0]);
    st->res0 = res0;
    st->res = res0;
    st->rz_old = out[1];
    st->iter = 0;
    st->converged = 0;
    st->stop = 0;
    if (hist) hist[0] = res0;
    if ((abs_tol ? res0 : res0 / (res0 + 1e-50)) < tol) {
        st->converged = 1;
        st->stop = 1;
    } else if (max_iters <= 1) {
        st->stop = 1;
    }
}

template<int MODE>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_pap(int64_t n, int64_t nb, Cols C, const double* __restrict__ val,
                                                        const double* __restrict__ p, double* __restrict__ Ap,
                                                        const CGState* st, double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t r = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (r < n) {
        const double y = row_dot_f64<MODE>(r, C, val, p);
        Ap[r] = y;
        v[0] = p[r] * y;
    }
    block_partials<1>(v, lb, nb, partials);
}

// k_cg_spmv_pap over the stencil copy (all lanes run the row product)
template<int NR>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_pap_st(int64_t n, int64_t nb, StCols S, const double* __restrict__ val,
                                                           const double* __restrict__ p, double* __restrict__ Ap,
                                                           const CGState* st, double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t r = lb * kBlock + threadIdx.x;
    const double y = row_dot_st<NR>(r, n, S, val, p);
    double v[1] = {0.0};
    if (r < n) {
        Ap[r] = y;
        v[0] = p[r] * y;
    }
    block_partials<1>(v, lb, nb, partials);
}

// block sums of p.Ap for an Ap formed by another kernel (a matrix-free apply)
__global__ __launch_bounds__(kBlock) void k_cg_pap(int64_t n, int64_t nb, const double* __restrict__ p,
                                                   const double* __restrict__ Ap, const CGState* st,
                                                   double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t r = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (r < n) v[0] = p[r] * Ap[r];
    block_partials<1>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kFold) void k_cg_alpha(int64_t nb, const double* __restrict__ partials, CGState* st) {
    if (st->stop) return;
    const double rz = st->rz_old;
    double out[1];
    fold_partials<1>(nb, partials, out);
    if (threadIdx.x == 0) st->alpha = rz / out[0];
}

__global__ __launch_bounds__(kBlock) void k_cg_update_r(int64_t n, int64_t nb, double* __restrict__ x,
                                                        const double* __restrict__ p, const double* __restrict__ Ap,
                                                        double* __restrict__ r, const double* __restrict__ invd,
                                                        const CGState* st, double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t i = lb * kBlock + threadIdx.x;
    const double alpha = st->alpha;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        x[i] += alpha * p[i];
        const double ri = r[i] + -alpha * Ap[i];
        r[i] = ri;
        const double zi = invd ? ri * invd[i] : ri;
        v[0] = ri * ri;
        v[1] = ri * zi;
    }
    block_partials<2>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kFold) void k_cg_beta(int64_t nb, const double* __restrict__ partials, CGState* st,
                                                   double tol, int32_t abs_tol, int64_t max_iters, double* hist) {
    if (st->stop) return;
    const int64_t k = st->iter + 1;
    const double res0 = st->res0, rz_old = st->rz_old;
    double out[2];
    fold_partials<2>(nb, partials, out);
    if (threadIdx.x != 0) return;
    st->iter = k;
    const double res = sqrt(out[0]);
    st->res = res;
    if (hist) hist[k] = res;
    if ((abs_tol ? res : res / (res0 + 1e-50)) < tol) {
        st->converged = 1;
        st->stop = 1;
    } else {
        st->beta = out[1] / rz_old;
        st->rz_old = out[1];
        if (k >= max_iters - 1) st->stop = 1;
    }
}

__global__ __launch_bounds__(kBlock) void k_cg_update_p(int64_t n, int64_t nb, double* __restrict__ p,
                                                        const double* __restrict__ r, const double* __restrict__ invd,
                                                        const CGState* st) {
    if (st->stop) return;
    const int64_t i = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    const double beta = st->beta;
    if (i < n) {
        const double zi = invd ? r[i] * invd[i] : r[i];
        p[i] = beta * p[i] + 1.0 * zi;
    }
}

}  // namespace rsd

using namespace rsd;

// ------------------------------------------------------------------------
// host-side launchers used by the drivers
// ------------------------------------------------------------------------
namespace rsd {

void launch_spmv_f64(rsd_ctx* ctx, const rsd_matrix* A, const double* x, double* y) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = nblocks(n);
    if (!n) return;
    RSD_LAUNCH_COLS(S.mode(), k_spmv_f64, dim3((unsigned)nb), dim3(kBlock), ctx->stream, n, nb, S.cols(), A->val.p,
                    x, y);
}

void reserve_stencil(rsd_ctx* ctx, const rsd_matrix* A) {
    const Sell& S = A->pb->sell;
    if (!S.has_st || A->val_st.n == S.st_slots()) return;
    A->val_st.alloc(S.st_slots());
    RSD_HIP(hipMemsetAsync(A->val_st.p, 0, sizeof(double) * (size_t)A->val_st.n, ctx->stream));
}

bool stencil_ready(const rsd_matrix* A) {
    const Sell& S = A->pb->sell;
    return S.has_st && S.n_rows > 0 && A->val_st.n == S.st_slots();
}

void launch_to_stencil(rsd_ctx* ctx, const rsd_matrix* A) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows;
    const dim3 grid((unsigned)nblocks(n)), block(kBlock);
    if (S.st_np == 27)
        hipLaunchKernelGGL(k_to_st<27>, grid, block, 0, ctx->stream, n, S.slice_ptr.p, S.st_mask.p, A->val.p, A->val_st.p);
    else
        hipLaunchKernelGGL(k_to_st<9>, grid, block, 0, ctx->stream, n, S.slice_ptr.p, S.st_mask.p, A->val.p, A->val_st.p);
}

void launch_spmv_stencil(rsd_ctx* ctx, const rsd_matrix* A, const double* x, double* y) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = nblocks(n);
    const dim3 grid((unsigned)nb), block(kBlock);
    if (S.st_np == 27)
        hipLaunchKernelGGL(k_spmv_st<9>, grid, block, 0, ctx->stream, n, nb, S.st_cols(), A->val_st.p, x, y);
    else
        hipLaunchKernelGGL(k_spmv_st<3>, grid, block, 0, ctx->stream, n, nb, S.st_cols(), A->val_st.p, x, y);
}

void launch_inv_diag(rsd_ctx* ctx, const rsd_matrix* A, double* inv) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows;
    if (!n) return;
    hipLaunchKernelGGL(k_inv_diag, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, S.diag.p, A->val.p, inv);
}

void launch_dot(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out, bool take_sqrt) {
    const int64_t nb = std::max<int64_t>(1, nblocks(n));
    double* part = ctx->partials(nb);
    hipLaunchKernelGGL(k_dot, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, n, nb, x, y, part);
    hipLaunchKernelGGL(k_dot_fold, dim3(1), dim3(kFold), 0, ctx->stream, nb, part, d_out, (int32_t)take_sqrt);
}

void launch_axpy(rsd_ctx* ctx, int64_t n, double a, const double* x, double* y) {
    if (n) hipLaunchKernelGGL(k_axpy, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, a, x, y);
}
void launch_axpby(rsd_ctx* ctx, int64_t n, double a, double b, const double* x, double* y) {
    if (n) hipLaunchKernelGGL(k_axpby, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, a, b, x, y);
}
void launch_scal(rsd_ctx* ctx, int64_t n, double a, double* x) {
    if (n) hipLaunchKernelGGL(k_scal, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, a, x);
}
void launch_copy(rsd_ctx* ctx, int64_t n, const double* x, double* y) {
    if (n) hipLaunchKernelGGL(k_copy, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, x, y);
}
void launch_fill(rsd_ctx* ctx, int64_t n, double* p, double v) {
    if (n) hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, p, v);
}

void launch_fill_f32(rsd_ctx* ctx, int64_t n, float* p, float v) {
    if (n) hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, p, v);
}

CGWork cg_workspace(rsd_ctx* ctx, int64_t n, int64_t max_iters, bool want_hist) {
    CGWork w;
    w.r = (double*)ctx->ws("cg.r", sizeof(double) * (size_t)std::max<int64_t>(n, 1));
    w.p = (double*)ctx->ws("cg.p", sizeof(double) * (size_t)std::max<int64_t>(n, 1));
    w.Ap = (double*)ctx->ws("cg.Ap", sizeof(double) * (size_t)std::max<int64_t>(n, 1));
    w.invd = (double*)ctx->ws("cg.invd", sizeof(double) * (size_t)std::max<int64_t>(n, 1));
    w.st = (CGState*)ctx->ws("cg.state", sizeof(CGState));
    w.hist = want_hist ? (double*)ctx->ws("cg.hist", sizeof(double) * (size_t)(max_iters + 1)) : nullptr;
    return w;
}

void cg_launch_init(rsd_ctx* ctx, const rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                    int32_t precond, const CGWork& w) {
    const int64_t n = A->pb->sell.n_rows, nb = std::max<int64_t>(1, nblocks(n));
    double* part = ctx->partials(3 * nb);
    if (precond == RSD_PC_JACOBI) launch_inv_diag(ctx, A, w.invd);
    if (stencil_ready(A)) {
        // the iterations read the stencil copy: refresh it from the values
        launch_to_stencil(ctx, A);
        launch_spmv_stencil(ctx, A, x, w.Ap);
    } else {
        launch_spmv_f64(ctx, A, x, w.Ap);
    }
    hipLaunchKernelGGL(k_cg_init, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, n, nb, b, w.Ap, w.r, w.p,
                       precond == RSD_PC_JACOBI ? w.invd : nullptr, part);
    hipLaunchKernelGGL(k_cg_init_fold, dim3(1), dim3(kFold), 0, ctx->stream, nb, part, w.st, tol, w.abs_tol, max_iters,
                       w.hist);
}

void cg_launch_iter(rsd_ctx* ctx, const rsd_matrix* A, double* x, double tol, int64_t max_iters, int32_t precond,
                    const CGWork& w) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = std::max<int64_t>(1, nblocks(n));
    double* part = ctx->partials(3 * nb);
    const double* invd = precond == RSD_PC_JACOBI ? w.invd : nullptr;
    hipStream_t s = ctx->stream;
    if (stencil_ready(A)) {
        if (S.st_np == 27)
            hipLaunchKernelGGL(k_cg_spmv_pap_st<9>, dim3((unsigned)nb), dim3(kBlock), 0, s, n, nb, S.st_cols(),
                               A->val_st.p, w.p, w.Ap, w.st, part);
        else
            hipLaunchKernelGGL(k_cg_spmv_pap_st<3>, dim3((unsigned)nb), dim3(kBlock), 0, s, n, nb, S.st_cols(),
                               A->val_st.p, w.p, w.Ap, w.st, part);
    } else {
        RSD_LAUNCH_COLS(S.mode(), k_cg_spmv_pap, dim3((unsigned)nb), dim3(kBlock), s, n, nb, S.cols(), A->val.p, w.p,
                        w.Ap, w.st, part);
    }
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kFold), 0, s, nb, part, w.st);
    hipLaunchKernelGGL(k_cg_update_r, dim3((unsigned)nb), dim3(kBlock), 0, s, n, nb, x, w.p, w.Ap, w.r, invd, w.st,
                       part + nb);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(kFold), 0, s, nb, part + nb, w.st, tol, w.abs_tol, max_iters, w.hist);
    hipLaunchKernelGGL(k_cg_update_p, dim3((unsigned)nb), dim3(kBlock), 0, s, n, nb, w.p, w.r, invd, w.st);
}

void cg_launch_init_ax(rsd_ctx* ctx, int64_t n, const double* b, double tol, int64_t max_iters, int32_t precond,
                       const CGWork& w) {
    const int64_t nb = std::max<int64_t>(1, nblocks(n));
    double* part = ctx->partials(3 * nb);
    hipLaunchKernelGGL(k_cg_init, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, n, nb, b, w.Ap, w.r, w.p,
                       precond == RSD_PC_JACOBI ? w.invd : nullptr, part);
    hipLaunchKernelGGL(k_cg_init_fold, dim3(1), dim3(kFold), 0, ctx->stream, nb, part, w.st, tol, w.abs_tol, max_iters,
                       w.hist);
}

void cg_launch_iter_ap(rsd_ctx* ctx, int64_t n, double* x, double tol, int64_t max_iters, int32_t precond,
                       const CGWork& w) {
    const int64_t nb = std::max<int64_t>(1, nblocks(n));
    double* part = ctx->partials(3 * nb);
    const double* invd = precond == RSD_PC_JACOBI ? w.invd : nullptr;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_cg_pap, dim3((unsigned)nb), dim3(kBlock), 0, s, n, nb, w.p, w.Ap, w.st, part);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kFold), 0, s, nb, part, w.st);
    hipLaunchKernelGGL(k_cg_update_r, dim3((unsigned)nb), dim3(kBlock), 0, s, n, nb, x, w.p, w.Ap, w.r, invd, w.st,
                       part + nb);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(kFold), 0, s, nb, part + nb, w.st, tol, w.abs_tol, max_iters, w.hist);
    hipLaunchKernelGGL(k_cg_update_p, dim3((unsigned)nb), dim3(kBlock), 0, s, n, nb, w.p, w.r, invd, w.st);
}

void cg_report(const CGState& s, int64_t max_iters, rsd_report* rep) {
    if (!rep) return;
    rep->converged = s.converged;
    rep->initial_residual = s.res0;
    rep->final_residual = s.converged || s.iter > 0 ? s.res : s.res0;
    rep->iterations = s.converged ? s.iter : max_iters;
    rep->final_residual_rel = s.converged && s.iter == 0 ? 1.0 : rep->final_residual / (s.res0 + 1e-50);
}

// Run a CG solve to completion with chunked launches; returns the final state.
CGState cg_run(rsd_ctx* ctx, const rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
               int32_t precond, const CGWork& w) {
    reserve_stencil(ctx, A);
    cg_launch_init(ctx, A, b, x, tol, max_iters, precond, w);
    CGState h{};
    CGState* pinned = nullptr;
    RSD_HIP(hipHostMalloc((void**)&pinned, sizeof(CGState), hipHostMallocDefault));
    int64_t chunk = 4, done_iters = 0;
    for (;;) {
        RSD_HIP(hipMemcpyAsync(pinned, w.st, sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
        RSD_HIP(hipStreamSynchronize(ctx->stream));
        h = *pinned;
        if (h.stop) break;
        RSD_CHECK(done_iters <= max_iters + 2, "CG did not stop within max_iters launches");
        const int64_t c = std::min<int64_t>(chunk, std::max<int64_t>(1, max_iters - done_iters));
        for (int64_t k = 0; k < c; ++k) cg_launch_iter(ctx, A, x, tol, max_iters, precond, w);
        RSD_HIP(hipGetLastError());
        done_iters += c;
        chunk = std::min<int64_t>(chunk * 2, 64);
    }
    (void)hipHostFree(pinned);
    return h;
}

// ------------------------------------------------------------------------
// The dot product for callers that cannot allocate: the block partials live
// in a file-scope device array instead of the context's (grown on demand)
// partials area.  Each device holds its own copy of the array; its address
// there is looked up once, with that device current, and kept.
// ------------------------------------------------------------------------
constexpr int64_t kScratchPartials = 65536;  // 512 KB: vectors of up to 2^24 entries
static __device__ __align__(256) double scratch[kScratchPartials];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kScratchDevices] = {};

static double* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    RSD_CHECK(device >= 0 && device < kScratchDevices, "rsd_dot_f64_noalloc: device index out of range");
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        RSD_HIP(hipGetDevice(&cur));
        if (cur != device) RSD_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        RSD_HIP(e);
        RSD_CHECK(p != nullptr, "rsd_dot_f64_noalloc: no device address for the scratch array");
        g_scratch_ptr[device] = (double*)p;
    }
    return g_scratch_ptr[device];
}

}  // namespace rsd

// ------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------
extern "C" {

int rsd_dot_f64_noalloc(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out) {
    return guard([&] {
        RSD_CHECK(ctx && n >= 0 && (n == 0 || (x && y)) && d_out, "rsd_dot_f64_noalloc: bad argument");
        const int64_t nb = std::max<int64_t>(1, nblocks(n));
        RSD_CHECK(nb <= kScratchPartials, "rsd_dot_f64_noalloc: vectors longer than 2^24 entries (the scratch array "
                                          "holds 65536 block partials of 256 entries)");
        double* part = scratch_ptr(ctx->device);
        // every block writes its partial and the fold reads exactly those:
        // nothing in the array needs clearing between calls
        hipLaunchKernelGGL(k_dot, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, n, nb, x, y, part);
        hipLaunchKernelGGL(k_dot_fold, dim3(1), dim3(kFold), 0, ctx->stream, nb, part, d_out, 0);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_fill_f64(rsd_ctx* ctx, double* p, int64_t n, double v) {
    return guard([&] {
        RSD_CHECK(ctx && (p || n == 0) && n >= 0, "rsd_fill_f64: bad argument");
        launch_fill(ctx, n, p, v);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_fill_f32(rsd_ctx* ctx, float* p, int64_t n, float v) {
    return guard([&] {
        RSD_CHECK(ctx && (p || n == 0) && n >= 0, "rsd_fill_f32: bad argument");
        if (n) hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, p, v);
        RSD_HIP(hipGetLastError());
    });
}

static void finish_scalar(rsd_ctx* ctx, double* d_out, double* h_out) {
    if (h_out) {
        RSD_HIP(hipMemcpyAsync(h_out, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        RSD_HIP(hipStreamSynchronize(ctx->stream));
    }
}

int rsd_dot_f64(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out, double* h_out) {
    return guard([&] {
        RSD_CHECK(ctx && n >= 0 && (n == 0 || (x && y)), "rsd_dot_f64: bad argument");
        double* o = d_out ? d_out : ctx->d_scalars;
        launch_dot(ctx, n, x, y, o, false);
        RSD_HIP(hipGetLastError());
        finish_scalar(ctx, o, h_out);
    });
}

int rsd_norm2_f64(rsd_ctx* ctx, int64_t n, const double* x, double* d_out, double* h_out) {
    return guard([&] {
        RSD_CHECK(ctx && n >= 0 && (n == 0 || x), "rsd_norm2_f64: bad argument");
        double* o = d_out ? d_out : ctx->d_scalars;
        launch_dot(ctx, n, x, x, o, true);
        RSD_HIP(hipGetLastError());
        finish_scalar(ctx, o, h_out);
    });
}

int rsd_axpy_f64(rsd_ctx* ctx, int64_t n, double a, const double* x, double* y) {
    return guard([&] { RSD_CHECK(ctx && n >= 0, "bad argument"); launch_axpy(ctx, n, a, x, y); RSD_HIP(hipGetLastError()); });
}
int rsd_axpby_f64(rsd_ctx* ctx, int64_t n, double a, double b, const double* x, double* y) {
    return guard([&] { RSD_CHECK(ctx && n >= 0, "bad argument"); launch_axpby(ctx, n, a, b, x, y); RSD_HIP(hipGetLastError()); });
}
int rsd_scal_f64(rsd_ctx* ctx, int64_t n, double a, double* x) {
    return guard([&] { RSD_CHECK(ctx && n >= 0, "bad argument"); launch_scal(ctx, n, a, x); RSD_HIP(hipGetLastError()); });
}
int rsd_copy_f64(rsd_ctx* ctx, int64_t n, const double* x, double* y) {
    return guard([&] { RSD_CHECK(ctx && n >= 0, "bad argument"); launch_copy(ctx, n, x, y); RSD_HIP(hipGetLastError()); });
}
int rsd_mul_f64(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* z) {
    return guard([&] {
        RSD_CHECK(ctx && n >= 0, "bad argument");
        if (n) hipLaunchKernelGGL(k_mul, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, x, y, z);
        RSD_HIP(hipGetLastError());
    });
}

static void dot_f32(rsd_ctx* ctx, int64_t n, const float* x, const float* y, double* d_out, bool take_sqrt) {
    const int64_t nb = std::max<int64_t>(1, nblocks(n));
    double* part = ctx->partials(nb);
    hipLaunchKernelGGL(k_dot_f32, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, n, nb, x, y, part);
    hipLaunchKernelGGL(k_dot_fold, dim3(1), dim3(kFold), 0, ctx->stream, nb, part, d_out, (int32_t)take_sqrt);
}

int rsd_dot_f32(rsd_ctx* ctx, int64_t n, const float* x, const float* y, double* d_out, double* h_out) {
    return guard([&] {
        RSD_CHECK(ctx && n >= 0 && (n == 0 || (x && y)), "rsd_dot_f32: bad argument");
        double* o = d_out ? d_out : ctx->d_scalars;
        dot_f32(ctx, n, x, y, o, false);
        RSD_HIP(hipGetLastError());
        finish_scalar(ctx, o, h_out);
    });
}

int rsd_norm2_f32(rsd_ctx* ctx, int64_t n, const float* x, double* d_out, double* h_out) {
    return guard([&] {
        RSD_CHECK(ctx && n >= 0 && (n == 0 || x), "rsd_norm2_f32: bad argument");
        double* o = d_out ? d_out : ctx->d_scalars;
        dot_f32(ctx, n, x, x, o, true);
        RSD_HIP(hipGetLastError());
        finish_scalar(ctx, o, h_out);
    });
}

#define RSD_F32_LAUNCH(kern, n, ...)                                                                              \
    do {                                                                                                          \
        if (n) hipLaunchKernelGGL(kern, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, __VA_ARGS__); \
        RSD_HIP(hipGetLastError());                                                                               \
    } while (0)

int rsd_axpy_f32(rsd_ctx* ctx, int64_t n, float a, const float* x, float* y) {
    return guard([&] { RSD_CHECK(ctx && n >= 0 && (n == 0 || (x && y)), "rsd_axpy_f32: bad argument");
                       RSD_F32_LAUNCH(k_axpy_f32, n, a, x, y); });
}
int rsd_axpby_f32(rsd_ctx* ctx, int64_t n, float a, float b, const float* x, float* y) {
    return guard([&] { RSD_CHECK(ctx && n >= 0 && (n == 0 || (x && y)), "rsd_axpby_f32: bad argument");
                       RSD_F32_LAUNCH(k_axpby_f32, n, a, b, x, y); });
}
int rsd_scal_f32(rsd_ctx* ctx, int64_t n, float a, float* x) {
    return guard([&] { RSD_CHECK(ctx && n >= 0 && (n == 0 || x), "rsd_scal_f32: bad argument");
                       RSD_F32_LAUNCH(k_scal_f32, n, a, x); });
}
int rsd_copy_f32(rsd_ctx* ctx, int64_t n, const float* x, float* y) {
    return guard([&] { RSD_CHECK(ctx && n >= 0 && (n == 0 || (x && y)), "rsd_copy_f32: bad argument");
                       RSD_F32_LAUNCH(k_copy_f32, n, x, y); });
}
int rsd_mul_f32(rsd_ctx* ctx, int64_t n, const float* x, const float* y, float* z) {
    return guard([&] { RSD_CHECK(ctx && n >= 0 && (n == 0 || (x && y && z)), "rsd_mul_f32: bad argument");
                       RSD_F32_LAUNCH(k_mul_f32, n, x, y, z); });
}

int rsd_spmv_f64(rsd_ctx* ctx, rsd_matrix* A, const double* x, double* y) {
    return guard([&] {
        RSD_CHECK(ctx && A && x && y, "rsd_spmv_f64: null argument");
        RSD_LIVE_MATRIX(A);
        launch_spmv_f64(ctx, A, x, y);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_problem_layout_stencil(rsd_problem* pb, int64_t* out) {
    return guard([&] {
        RSD_CHECK(pb && out, "null argument");
        const Sell& S = pb->sell;
        out[0] = S.has_st ? S.st_np : 0;
        out[1] = S.has_st ? S.st_slots() : 0;
    });
}

int rsd_problem_layout_union(rsd_problem* pb, int64_t* out) {
    return guard([&] {
        RSD_CHECK(pb && out, "null argument");
        const Sell& S = pb->sell;
        out[0] = S.has_su ? 1 : 0;
        out[1] = S.has_su ? S.su_slots : 0;
    });
}

int rsd_matrix_to_stencil(rsd_ctx* ctx, rsd_matrix* A) {
    return guard([&] {
        RSD_CHECK(ctx && A, "null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(A->pb->sell.has_st, "rsd_matrix_to_stencil: the matrix's layout has no stencil copy");
        reserve_stencil(ctx, A);
        if (A->pb->sell.n_rows) launch_to_stencil(ctx, A);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_spmv_f64_stencil(rsd_ctx* ctx, rsd_matrix* A, const double* x, double* y) {
    return guard([&] {
        RSD_CHECK(ctx && A && x && y, "rsd_spmv_f64_stencil: null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(stencil_ready(A), "rsd_spmv_f64_stencil: call rsd_matrix_to_stencil first");
        launch_spmv_stencil(ctx, A, x, y);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_jacobi_inv_diag(rsd_ctx* ctx, rsd_matrix* A, double* inv) {
    return guard([&] {
        RSD_CHECK(ctx && A && inv, "rsd_jacobi_inv_diag: null argument");
        RSD_LIVE_MATRIX(A);
        launch_inv_diag(ctx, A, inv);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_cg_solve_f64(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int32_t tol_type,
                     int64_t max_iters, int32_t precond, rsd_report* rep, double* hist) {
    return guard([&] {
        RSD_CHECK(ctx && A && b && x, "rsd_cg_solve_f64: null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        RSD_CHECK(tol_type == RSD_TOL_RELATIVE || tol_type == RSD_TOL_ABSOLUTE, "unknown tolerance type");
        RSD_CHECK(max_iters >= 0, "max_iters must be >= 0");
        const int64_t n = A->pb->sell.n_rows;
        CGWork w = cg_workspace(ctx, n, max_iters, hist != nullptr);
        w.abs_tol = tol_type == RSD_TOL_ABSOLUTE;
        CGState s = cg_run(ctx, A, b, x, tol, max_iters, precond, w);
        cg_report(s, max_iters, rep);
        if (hist) {
            const int64_t cnt = std::min<int64_t>(s.iter + 1, max_iters + 1);
            RSD_HIP(hipMemcpyAsync(hist, w.hist, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost, ctx->stream));
            RSD_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

}  // extern "C"
