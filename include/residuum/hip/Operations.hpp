// residuum/hip/Operations.hpp -- operations::* for Hip containers.
//
// The reference declares its BLAS-1 and matrix operations as free function
// templates (linalg/operations/VectorOps.hpp:11-37, MatrixOps.hpp:11-15)
// whose bodies come from backend/cpu/*.tpp.  These non-template overloads
// for Vector<double, Hip>, Vector<float, Hip> and CSRMatrix<double, Hip> are preferred over
// those templates by overload resolution, and run on the device:
//   dot / norm      fixed-tree reductions (bitwise reproducible; the sum
//                   order differs from the CPU loop, so results agree with
//                   it to rounding, not bitwise)
//   axpy ... copy   elementwise, bitwise the CPU loop
//   matvec / lump   row sums left to right, bitwise the CPU loop
// A template of the reference that calls operations::* on Hip vectors must
// see these declarations: include this header before the reference's
// header that defines that template.  A qualified call such as
// operations::copy(b, W.r) in a template body binds to the overloads
// declared before that body; one included earlier would compile and then
// run the CPU loops on device memory, so that order is refused here.
#pragma once

#if defined(RESIDUUM_LINALG_SOLVER_ITERATIVE_CG_SOLVER_HPP) ||          \
    defined(RESIDUUM_LINALG_SOLVER_ITERATIVE_GMRES_SOLVER_HPP) ||       \
    defined(RESIDUUM_LINALG_OPERATOR_CSROPERATOR_HPP) ||                \
    defined(RESIDUUM_LINALG_SOLVER_PRECONDITIONER_IDENTITY_HPP) ||      \
    defined(RESIDUUM_LINALG_SOLVER_PRECONDITIONER_JACOBI_HPP) ||        \
    defined(RESIDUUM_SOLVER_STAGE_BACKWARDEULERSTAGE_HPP) ||            \
    defined(RESIDUUM_SOLVER_STAGE_STEADYSTAGE_HPP) ||                   \
    defined(RESIDUUM_APPLICATION_HEATEQ_PROBLEM_HEATPROBLEM_HPP)
#error "residuum/hip/Operations.hpp must be included before the reference headers whose templates call operations::*"
#endif

#include <cassert>

#include "residuum/hip/DOFNorms.hpp"
#include "linalg/operations/MatrixOps.hpp"
#include "linalg/operations/VectorOps.hpp"
#include "linalg/types/CSRMatrix.hpp"
#include "linalg/types/Vector.hpp"
#include "residuum/hip/Backend.hpp"

namespace residuum {
namespace hip {
using HipVector = linalg::types::Vector<double, linalg::types::backend::Hip>;
using HipVectorF = linalg::types::Vector<float, linalg::types::backend::Hip>;
using HipCSRMatrix = linalg::types::CSRMatrix<double, linalg::types::backend::Hip>;
static_assert(sizeof(Index) == sizeof(uint64_t), "the device CSR uses 64-bit indices like the reference's size_t");
}  // namespace hip

namespace linalg {
namespace operations {

// VectorOps.tpp:7-19
inline double dot(const hip::HipVector& a, const hip::HipVector& b) {
    assert(a.size() == b.size());
    double r = 0.0;
    hip::check(rsd_dot_f64(hip::Context::get(), (int64_t)a.size(), a.data(), b.data(), nullptr, &r));
    return r;
}

// VectorOps.tpp:22-31
inline void axpy(const double alpha, const hip::HipVector& x, hip::HipVector& y) {
    assert(x.size() == y.size());
    hip::check(rsd_axpy_f64(hip::Context::get(), (int64_t)x.size(), alpha, x.data(), y.data()));
}

// VectorOps.tpp:34-43
inline void axpby(const double alpha, const double beta, const hip::HipVector& x, hip::HipVector& y) {
    assert(x.size() == y.size());
    hip::check(rsd_axpby_f64(hip::Context::get(), (int64_t)x.size(), alpha, beta, x.data(), y.data()));
}

// VectorOps.tpp:46-53
inline void scal(const double alpha, hip::HipVector& x) {
    hip::check(rsd_scal_f64(hip::Context::get(), (int64_t)x.size(), alpha, x.data()));
}

// VectorOps.tpp:56-65
inline void copy(const hip::HipVector& x, hip::HipVector& y) {
    assert(x.size() == y.size());
    hip::check(rsd_copy_f64(hip::Context::get(), (int64_t)x.size(), x.data(), y.data()));
}

// VectorOps.tpp:68-78
inline void multiply(const hip::HipVector& x, const hip::HipVector& y, hip::HipVector& z) {
    assert(x.size() == y.size() && y.size() == z.size());
    hip::check(rsd_mul_f64(hip::Context::get(), (int64_t)x.size(), x.data(), y.data(), z.data()));
}

// VectorOps.tpp:81-86
inline double norm(const hip::HipVector& x) {
    double r = 0.0;
    hip::check(rsd_norm2_f64(hip::Context::get(), (int64_t)x.size(), x.data(), nullptr, &r));
    return r;
}

// The same operations on float vectors: elementwise bitwise the CPU loop;
// dot and norm sum exact FP64 products (the CPU loop sums in float) and
// return the result rounded to float.
inline float dot(const hip::HipVectorF& a, const hip::HipVectorF& b) {
    assert(a.size() == b.size());
    double r = 0.0;
    hip::check(rsd_dot_f32(hip::Context::get(), (int64_t)a.size(), a.data(), b.data(), nullptr, &r));
    return (float)r;
}

inline void axpy(const float alpha, const hip::HipVectorF& x, hip::HipVectorF& y) {
    assert(x.size() == y.size());
    hip::check(rsd_axpy_f32(hip::Context::get(), (int64_t)x.size(), alpha, x.data(), y.data()));
}

inline void axpby(const float alpha, const float beta, const hip::HipVectorF& x, hip::HipVectorF& y) {
    assert(x.size() == y.size());
    hip::check(rsd_axpby_f32(hip::Context::get(), (int64_t)x.size(), alpha, beta, x.data(), y.data()));
}

inline void scal(const float alpha, hip::HipVectorF& x) {
    hip::check(rsd_scal_f32(hip::Context::get(), (int64_t)x.size(), alpha, x.data()));
}

inline void copy(const hip::HipVectorF& x, hip::HipVectorF& y) {
    assert(x.size() == y.size());
    hip::check(rsd_copy_f32(hip::Context::get(), (int64_t)x.size(), x.data(), y.data()));
}

inline void multiply(const hip::HipVectorF& x, const hip::HipVectorF& y, hip::HipVectorF& z) {
    assert(x.size() == y.size() && y.size() == z.size());
    hip::check(rsd_mul_f32(hip::Context::get(), (int64_t)x.size(), x.data(), y.data(), z.data()));
}

inline float norm(const hip::HipVectorF& x) {
    double r = 0.0;
    hip::check(rsd_norm2_f32(hip::Context::get(), (int64_t)x.size(), x.data(), nullptr, &r));
    return (float)r;
}

// MatrixOps.tpp:5-23
inline void matvec(const hip::HipCSRMatrix& A, const hip::HipVector& x, hip::HipVector& y) {
    assert(A.nCols() == x.size());
    assert(A.nRows() == y.size());
    hip::check(rsd_csr_spmv_f64(hip::Context::get(), (int64_t)A.nRows(), reinterpret_cast<const uint64_t*>(A.rowPtr()),
                                reinterpret_cast<const uint64_t*>(A.colIdx()), A.data(), x.data(), y.data()));
}

// MatrixOps.tpp:26-44
inline void lump(const hip::HipCSRMatrix& A, hip::HipVector& diag) {
    assert(A.nRows() == A.nCols());
    assert(A.nRows() == diag.size());
    hip::check(rsd_csr_lump_f64(hip::Context::get(), (int64_t)A.nRows(), reinterpret_cast<const uint64_t*>(A.rowPtr()),
                                A.data(), diag.data()));
}

}  // namespace operations
}  // namespace linalg
}  // namespace residuum
