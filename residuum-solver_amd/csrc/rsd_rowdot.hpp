// rsd_rowdot.hpp -- the sliced-ELL row product shared by every SpMV kernel.
#pragma once

#include "rsd_internal.hpp"

namespace rsd {

// y_r = sum_j val * x[col] over row r, in CSR order, each product and sum
// rounded separately: the reference's left-to-right row sum
// (linalg/operations/backend/cpu/MatrixOps.tpp:5-23) bit for bit.
__device__ __forceinline__ double row_dot_f64(int64_t r, const int64_t* __restrict__ slice_ptr,
                                              const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                              const double* __restrict__ val, const double* __restrict__ x) {
    const int64_t base = slice_ptr[r >> 6] + (r & 63);
    const int len = rp[r + 1] - rp[r];
    double sum = 0.0;
    int j = 0;
    for (; j + 4 <= len; j += 4) {
        const int64_t s0 = base + (int64_t)j * kSlice;
        const int c0 = col[s0], c1 = col[s0 + kSlice], c2 = col[s0 + 2 * kSlice], c3 = col[s0 + 3 * kSlice];
        const double v0 = val[s0], v1 = val[s0 + kSlice], v2 = val[s0 + 2 * kSlice], v3 = val[s0 + 3 * kSlice];
        const double x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        sum += v0 * x0;
        sum += v1 * x1;
        sum += v2 * x2;
        sum += v3 * x3;
    }
    for (; j < len; ++j) {
        const int64_t s0 = base + (int64_t)j * kSlice;
        sum += val[s0] * x[col[s0]];
    }
    return sum;
}

}  // namespace rsd
