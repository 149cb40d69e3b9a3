// rsd_csr.hip -- operations on a raw device CSR in the reference's own
// layout (size_t row pointer and column indices, linalg/types/CSRMatrix.hpp),
// and the bridge from such a CSR to the sliced-ELL matrices the fused
// solvers run on.
//
// These serve the generic seam: a caller that keeps the reference's
// CSRMatrix<T, Hip> (values and indices in device memory) gets matvec,
// lump and diagonal extraction with the reference's arithmetic.  Row sums
// are sequential in CSR order, so the products are bitwise those of
// linalg/operations/backend/cpu/MatrixOps.tpp:5-44.
#include "rsd_linalg.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace rsd {

// y_i = sum_p val[p] * x[col[p]], p over row i in order (MatrixOps.tpp:5-23)
__global__ void k_csr_spmv(int64_t n, const uint64_t* __restrict__ rp, const uint64_t* __restrict__ ci,
                           const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double sum = 0.0;
    const uint64_t e = rp[i + 1];
    for (uint64_t p = rp[i]; p < e; ++p) sum += val[p] * x[ci[p]];
    y[i] = sum;
}

// d_i = sum_p val[p] over row i (MatrixOps.tpp:26-44)
__global__ void k_csr_lump(int64_t n, const uint64_t* __restrict__ rp, const double* __restrict__ val,
                           double* __restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double sum = 0.0;
    const uint64_t e = rp[i + 1];
    for (uint64_t p = rp[i]; p < e; ++p) sum += val[p];
    d[i] = sum;
}

// d_i = A(i, i) by binary search in the row (CSROperator.hpp:40-45,
// CSRMatrix::getDataIndex); a missing diagonal raises the error flag
__global__ void k_csr_diag(int64_t n, const uint64_t* __restrict__ rp, const uint64_t* __restrict__ ci,
                           const double* __restrict__ val, double* __restrict__ d, int32_t* missing) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t lo = rp[i], hi = rp[i + 1];
    const uint64_t want = (uint64_t)i;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        const uint64_t c = ci[mid];
        if (c == want) {
            d[i] = val[mid];
            return;
        }
        if (c < want) lo = mid + 1;
        else hi = mid;
    }
    missing[0] = 1;
}

// Jacobi::update (Jacobi.hpp:33-36): 1/d where |d| > 1e-300, else 1 (d and
// inv may be the same array)
__global__ void k_inv_rule(int64_t n, const double* d, double* inv) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double v = d[i];
    inv[i] = (fabs(v) > 1e-300) ? 1.0 / v : 1.0;
}

__global__ void k_clear_flag(int32_t* flag) { flag[0] = 0; }

// SELL values from CSR-order values
__global__ void k_csr_to_sell(int64_t nnz, const int32_t* __restrict__ slot, const double* __restrict__ csr,
                              double* __restrict__ sell) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < nnz) sell[slot[k]] = csr[k];
}

}  // namespace rsd

using namespace rsd;

extern "C" {

int rsd_csr_spmv_f64(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col_idx,
                     const double* val, const double* x, double* y) {
    return guard([&] {
        RSD_CHECK(ctx && n_rows >= 0, "rsd_csr_spmv_f64: bad argument");
        if (!n_rows) return;
        RSD_CHECK(row_ptr && col_idx && val && x && y, "rsd_csr_spmv_f64: null argument");
        hipLaunchKernelGGL(k_csr_spmv, dim3((unsigned)nblocks(n_rows)), dim3(kBlock), 0, ctx->stream, n_rows, row_ptr,
                           col_idx, val, x, y);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_csr_lump_f64(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const double* val, double* d) {
    return guard([&] {
        RSD_CHECK(ctx && n_rows >= 0, "rsd_csr_lump_f64: bad argument");
        if (!n_rows) return;
        RSD_CHECK(row_ptr && val && d, "rsd_csr_lump_f64: null argument");
        hipLaunchKernelGGL(k_csr_lump, dim3((unsigned)nblocks(n_rows)), dim3(kBlock), 0, ctx->stream, n_rows, row_ptr,
                           val, d);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_csr_diag_f64(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col_idx,
                     const double* val, double* d) {
    return guard([&] {
        RSD_CHECK(ctx && n_rows >= 0, "rsd_csr_diag_f64: bad argument");
        if (!n_rows) return;
        RSD_CHECK(row_ptr && col_idx && val && d, "rsd_csr_diag_f64: null argument");
        int32_t* flag = (int32_t*)ctx->ws("csr.missing", sizeof(int32_t));
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, ctx->stream, flag);
        hipLaunchKernelGGL(k_csr_diag, dim3((unsigned)nblocks(n_rows)), dim3(kBlock), 0, ctx->stream, n_rows, row_ptr,
                           col_idx, val, d, flag);
        int32_t h = 0;
        RSD_HIP(hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        RSD_HIP(hipStreamSynchronize(ctx->stream));
        RSD_CHECK(h == 0, "Index does not correspond to non-zero entry.");
    });
}

int rsd_inv_diag_f64(rsd_ctx* ctx, int64_t n, const double* d, double* inv) {
    return guard([&] {
        RSD_CHECK(ctx && n >= 0, "rsd_inv_diag_f64: bad argument");
        if (!n) return;
        RSD_CHECK(d && inv, "rsd_inv_diag_f64: null argument");
        hipLaunchKernelGGL(k_inv_rule, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, d, inv);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_matrix_from_csr(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col_idx,
                        rsd_matrix** out) {
    return guard([&] {
        RSD_CHECK(ctx && out && n_rows >= 0, "rsd_matrix_from_csr: bad argument");
        RSD_CHECK(n_rows == 0 || (row_ptr && col_idx), "rsd_matrix_from_csr: null argument");
        hipStream_t st = ctx->stream;
        std::vector<uint64_t> rp(n_rows + 1, 0);
        if (n_rows) RSD_HIP(hipMemcpyAsync(rp.data(), row_ptr, sizeof(uint64_t) * (n_rows + 1), hipMemcpyDeviceToHost, st));
        RSD_HIP(hipStreamSynchronize(st));
        RSD_CHECK(rp[0] == 0, "rsd_matrix_from_csr: row pointer does not start at 0");
        for (int64_t r = 0; r < n_rows; ++r) RSD_CHECK(rp[r + 1] >= rp[r], "rsd_matrix_from_csr: row pointer decreases");
        const uint64_t nnz = rp[n_rows];
        RSD_CHECK(nnz < (uint64_t)INT32_MAX, "rsd_matrix_from_csr: more than 2^31 - 1 nonzeros");
        std::vector<uint64_t> ci(nnz);
        if (nnz) RSD_HIP(hipMemcpyAsync(ci.data(), col_idx, sizeof(uint64_t) * nnz, hipMemcpyDeviceToHost, st));
        RSD_HIP(hipStreamSynchronize(st));
        auto pb = std::make_unique<rsd_problem>();
        pb->ctx = ctx;
        pb->n_free = n_rows;
        pb->nnz = (int64_t)nnz;
        pb->h_rp.assign(n_rows + 1, 0);
        pb->h_ci.resize(nnz);
        std::vector<int32_t> rowlen(n_rows);
        for (int64_t r = 0; r < n_rows; ++r) {
            pb->h_rp[r + 1] = (int32_t)rp[r + 1];
            rowlen[r] = (int32_t)(rp[r + 1] - rp[r]);
            for (uint64_t p = rp[r]; p < rp[r + 1]; ++p) {
                RSD_CHECK(ci[p] < (uint64_t)n_rows, "rsd_matrix_from_csr: column outside a square matrix");
                RSD_CHECK(p == rp[r] || ci[p] > ci[p - 1], "rsd_matrix_from_csr: columns not sorted and unique");
                pb->h_ci[p] = (int32_t)ci[p];
            }
        }
        build_sell(pb.get(), rowlen);
        std::vector<int32_t> slot(nnz);
        for (int64_t r = 0; r < n_rows; ++r)
            for (int64_t j = 0; j < rowlen[r]; ++j)
                slot[pb->h_rp[r] + j] = (int32_t)(pb->sell.h_slice_ptr[r / kSlice] + j * kSlice + (r % kSlice));
        pb->csr_slot.upload(slot.data(), (int64_t)nnz, st);
        auto m = std::make_unique<rsd_matrix>();
        m->pb = pb.get();
        m->ctx = ctx;
        m->val.alloc(pb->sell.slots);
        if (pb->sell.slots) RSD_HIP(hipMemsetAsync(m->val.p, 0, sizeof(double) * pb->sell.slots, st));
        m->own_pb = std::move(pb);
        RSD_HIP(hipStreamSynchronize(st));
        *out = m.release();
    });
}

int rsd_matrix_load_csr(rsd_ctx* ctx, rsd_matrix* A, const double* val) {
    return guard([&] {
        RSD_CHECK(ctx && A && val, "rsd_matrix_load_csr: null argument");
        RSD_LIVE_MATRIX(A);
        if (A->pb->csr_slot.n != A->pb->nnz) {  // a problem's matrix: slots from the sliced layout
            RSD_CHECK(ctx == A->pb->ctx, "matrix belongs to another context");
            launch_sell_csr(A, const_cast<double*>(val), false);
            return;
        }
        if (A->pb->nnz)
            hipLaunchKernelGGL(k_csr_to_sell, dim3((unsigned)nblocks(A->pb->nnz)), dim3(kBlock), 0, ctx->stream,
                               A->pb->nnz, A->pb->csr_slot.p, val, A->val.p);
        RSD_HIP(hipGetLastError());
    });
}

}  // extern "C"
