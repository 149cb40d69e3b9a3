// rsd_rowdot.hpp -- the sliced-ELL row product shared by every SpMV kernel.
#pragma once

#include "rsd_internal.hpp"

namespace rsd {

// Matrix values and column indices are read once per product: they are
// loaded non-temporally so the stream does not evict the x entries the
// neighbouring rows gather from L2.
template<typename T>
__device__ __forceinline__ T stream_load(const T* p) {
    return __builtin_nontemporal_load(p);
}

// y_r = sum_j val * x[col] over row r, in CSR order, each product and sum
// rounded separately: the reference's left-to-right row sum
// (linalg/operations/backend/cpu/MatrixOps.tpp:5-23) bit for bit.  Eight
// slots' loads are issued before their products are added in order.
__device__ __forceinline__ double row_dot_f64(int64_t r, const int64_t* __restrict__ slice_ptr,
                                              const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                              const double* __restrict__ val, const double* __restrict__ x) {
    constexpr int U = 8;
    const int64_t base = slice_ptr[r >> 6] + (r & 63);
    const int len = rp[r + 1] - rp[r];
    double sum = 0.0;
    int j = 0;
    for (; j + U <= len; j += U) {
        const int64_t s0 = base + (int64_t)j * kSlice;
        int c[U];
        double v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            c[k] = stream_load(col + s0 + k * kSlice);
            v[k] = stream_load(val + s0 + k * kSlice);
        }
        double xv[U];
#pragma unroll
        for (int k = 0; k < U; ++k) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 0; k < U; ++k) sum += v[k] * xv[k];
    }
    for (; j < len; ++j) {
        const int64_t s0 = base + (int64_t)j * kSlice;
        sum += stream_load(val + s0) * x[stream_load(col + s0)];
    }
    return sum;
}

}  // namespace rsd
