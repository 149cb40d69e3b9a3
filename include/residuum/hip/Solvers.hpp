// residuum/hip/Solvers.hpp -- the operator, preconditioners and linear
// solvers of the reference specialised for Hip containers.
//
//   op::CSROperator<CSRMatrix<double, Hip>>      linalg/operator/CSROperator.hpp:13-47
//   preconditioner::Jacobi<Vector<double, Hip>>  linalg/solver/preconditioner/Jacobi.hpp:14-49
//   preconditioner::Identity<Vector<double, Hip>> .../Identity.hpp:11-27
//   iterative::cg::Solver<CSROperator<...Hip>, Vector<...Hip>, M, Logger>
//                                                linalg/solver/iterative/cg/Solver.hpp:21-36
//   iterative::gmres::Solver<...>                linalg/solver/iterative/gmres/Solver.hpp:22-37
//
// The solver specialisations run the fused device solvers of
// libresiduum_hip.so (every iteration's scalars stay on the device) on a
// sliced-ELL copy of the operator's matrix: its pattern is built once per
// CSR pattern (the row pointer / column arrays), its values are reloaded on
// every solve.  The Workspace argument is accepted for interface
// compatibility; the device solvers keep their own.  The logger sees the
// same calls as from the reference's loop: reset, one log() per iteration
// with the residual 2-norm (the single-field per-DOF norm), summary.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <memory>
#include <type_traits>
#include <vector>

// first: the reference's cg / gmres Solver bodies (Solver.tpp) call
// operations::* with qualified names, which bind to the overloads declared
// before them -- the Hip ones must precede those bodies, or the generic
// solvers (used over operators this header does not specialise, e.g.
// FEMOperator) would run the CPU loops on device memory
#include "residuum/hip/Operations.hpp"

#include "linalg/operator/CSROperator.hpp"
#include "linalg/operator/Operator.hpp"
#include "linalg/solver/base/SolverReport.hpp"
#include "linalg/solver/iterative/cg/Solver.hpp"
#include "linalg/solver/iterative/gmres/Solver.hpp"
#include "linalg/solver/preconditioner/Identity.hpp"
#include "linalg/solver/preconditioner/Jacobi.hpp"

namespace residuum {
namespace linalg {
namespace op {

template<>
class CSROperator<hip::HipCSRMatrix> {
public:
    using value_type = double;

    const hip::HipCSRMatrix& A;

    explicit CSROperator(const hip::HipCSRMatrix& mat) : A(mat) {}

    template<typename VectorT>
    void apply(const VectorT& x, VectorT& y) const {
        operations::matvec(A, x, y);
    }

    Index size() const { return A.nRows(); }

    // SpMV: one multiply + one add per nonzero
    Index flopsPerApply() const { return 2 * nnz(); }

    // CSROperator.hpp:40-45: the diagonal entry of every row; a missing one
    // throws std::out_of_range like CSRMatrix::getDataIndex (CSRMatrix.hpp:70)
    template<typename VectorT>
    void diagonal(VectorT& d) const {
        const int rc = rsd_csr_diag_f64(hip::Context::get(), (int64_t)A.nRows(),
                                        reinterpret_cast<const uint64_t*>(A.rowPtr()),
                                        reinterpret_cast<const uint64_t*>(A.colIdx()), A.data(), d.data());
        if (rc == RSD_ERR_ARG) throw std::out_of_range(rsd_last_error_string());
        hip::check(rc);
    }

    Index nnz() const {
        Index v = 0;
        if (A.nRows())
            linalg::types::backend::Hip::copy<Index>(&v, A.rowPtr() + A.nRows(), 1,
                                                     linalg::types::backend::CopyKind::DeviceToHost);
        return v;
    }
};

}  // namespace op

namespace solver {
namespace preconditioner {

template<>
class Jacobi<hip::HipVector> {
public:
    using DataType = double;

    Jacobi() = default;

    template<typename OperatorT>
    requires linalg::op::DiagonalExtractable<OperatorT, hip::HipVector>
    void update(const OperatorT& A) {
        if (!invDiag_) invDiag_ = std::make_unique<hip::HipVector>(A.size());
        A.diagonal(*invDiag_);
        hip::check(rsd_inv_diag_f64(hip::Context::get(), (int64_t)invDiag_->size(), invDiag_->data(),
                                    invDiag_->data()));
    }

    void apply(const hip::HipVector& r, hip::HipVector& z) { operations::multiply(r, *invDiag_, z); }

    Index flopsPerApply() const { return invDiag_ ? invDiag_->size() : Index(0); }

private:
    std::unique_ptr<hip::HipVector> invDiag_;
};

template<>
class Identity<hip::HipVector> {
public:
    void apply(const hip::HipVector& r, hip::HipVector& z) { operations::copy(r, z); }

    template<typename OperatorT>
    void update(const OperatorT&) {}

    Index flopsPerApply() const { return 0; }
};

}  // namespace preconditioner
}  // namespace solver
}  // namespace linalg

namespace hip {

// The solver-side copy of a CSR operator: libresiduum_hip.so's sliced-ELL
// matrix on the same pattern, rebuilt when the pattern arrays change.
class SolverMatrix {
public:
    SolverMatrix() = default;
    SolverMatrix(const SolverMatrix&) = delete;
    SolverMatrix& operator=(const SolverMatrix&) = delete;
    ~SolverMatrix() {
        if (m_) rsd_matrix_destroy(m_);
    }

    rsd_matrix* load(const HipCSRMatrix& A) {
        const void* rp = A.rowPtr();
        const void* ci = A.colIdx();
        if (!m_ || rp != rp_ || ci != ci_ || A.nRows() != n_) {
            if (m_) rsd_matrix_destroy(m_);
            m_ = nullptr;
            if (rsd_problem* pb = detail::pattern_problem(rp, ci))
                check(rsd_matrix_create(pb, &m_));  // the assembler's pattern: its sliced layout as it is
            else
                check(rsd_matrix_from_csr(Context::get(), (int64_t)A.nRows(), reinterpret_cast<const uint64_t*>(rp),
                                          reinterpret_cast<const uint64_t*>(ci), &m_));
            rp_ = rp;
            ci_ = ci;
            n_ = A.nRows();
        }
        check(rsd_matrix_load_csr(Context::get(), m_, A.data()));
        return m_;
    }

private:
    rsd_matrix* m_ = nullptr;
    const void* rp_ = nullptr;
    const void* ci_ = nullptr;
    Index n_ = 0;
};

template<typename PreconditionerT>
constexpr int32_t precond_code() {
    if constexpr (std::is_same_v<PreconditionerT, linalg::solver::preconditioner::Jacobi<HipVector>>)
        return RSD_PC_JACOBI;
    else if constexpr (std::is_same_v<PreconditionerT, linalg::solver::preconditioner::Identity<HipVector>>)
        return RSD_PC_IDENTITY;
    else
        static_assert(sizeof(PreconditionerT) == 0, "the device solvers take Jacobi or Identity preconditioning");
}

template<typename ReportT>
void fill_report(ReportT& report, const rsd_report& r) {
    report.converged = r.converged != 0;
    report.iterations = (Index)r.iterations;
    report.initialResidual = r.initial_residual;
    report.finalResidual = r.final_residual;
    report.finalResidualRel = r.final_residual_rel;
    if (report.converged) report.perFieldResidual = {r.final_residual};
}

}  // namespace hip

namespace linalg {
namespace solver {
namespace iterative {
namespace cg {

template<typename PreconditionerT, typename LoggerT>
class Solver<op::CSROperator<hip::HipCSRMatrix>, hip::HipVector, PreconditionerT, LoggerT> {
public:
    using Config = residuum::linalg::solver::iterative::cg::Config<hip::HipVector>;
    using Workspace = residuum::linalg::solver::iterative::cg::Workspace<hip::HipVector>;

    Solver(const Config& cfg) : config_(cfg), mat_(std::make_shared<hip::SolverMatrix>()) {}

    // Solver.tpp:15-119
    bool solve(solver::SolverReport<hip::HipVector>& report, LoggerT& logger, Workspace&, PreconditionerT& M,
               const op::CSROperator<hip::HipCSRMatrix>& A, const hip::HipVector& b, hip::HipVector& x) {
        logger.reset();
        M.update(A);
        const double flopsPerIter = (double)A.flopsPerApply() + (double)M.flopsPerApply() + 10.0 * (double)x.size();
        rsd_matrix* S = mat_->load(A.A);
        const int64_t maxIters = (int64_t)config_.maxIters;
        std::vector<double> hist((size_t)maxIters + 1, 0.0);
        rsd_report r{};
        hip::check(rsd_cg_solve_f64(hip::Context::get(), S, b.data(), x.data(), config_.tol,
                                    config_.tolType == ToleranceType::Relative ? RSD_TOL_RELATIVE : RSD_TOL_ABSOLUTE,
                                    maxIters, hip::precond_code<PreconditionerT>(), &r, hist.data()));
        const int64_t last = r.converged ? r.iterations : std::max<int64_t>(maxIters - 1, 0);
        for (int64_t k = 0; k <= last && k <= maxIters; ++k)
            logger.log((Index)k, std::vector<double>{hist[(size_t)k]}, k == 0 ? 0.0 : flopsPerIter);
        hip::fill_report(report, r);
        logger.summary(report.converged);
        return report.converged;
    }

private:
    Config config_;
    std::shared_ptr<hip::SolverMatrix> mat_;
};

}  // namespace cg

namespace gmres {

template<typename PreconditionerT, typename LoggerT>
class Solver<op::CSROperator<hip::HipCSRMatrix>, hip::HipVector, PreconditionerT, LoggerT> {
public:
    using Config = residuum::linalg::solver::iterative::gmres::Config<hip::HipVector>;
    using Workspace = residuum::linalg::solver::iterative::gmres::Workspace<hip::HipVector>;

    Solver(const Config& cfg) : config_(cfg), mat_(std::make_shared<hip::SolverMatrix>()) {}

    // gmres/Solver.tpp:5-170; the logger sees the initial and the final residual
    bool solve(solver::SolverReport<hip::HipVector>& report, LoggerT& logger, Workspace&, PreconditionerT& M,
               const op::CSROperator<hip::HipCSRMatrix>& A, const hip::HipVector& b, hip::HipVector& x) {
        logger.reset();
        M.update(A);
        rsd_matrix* S = mat_->load(A.A);
        rsd_report r{};
        hip::check(rsd_gmres_solve_f64(hip::Context::get(), S, b.data(), x.data(), config_.tol,
                                       config_.tolType == ToleranceType::Relative ? RSD_TOL_RELATIVE : RSD_TOL_ABSOLUTE,
                                       (int64_t)config_.maxIters, (int32_t)config_.krylovDim,
                                       hip::precond_code<PreconditionerT>(), &r));
        logger.log(Index(0), std::vector<double>{r.initial_residual}, 0.0);
        if (r.iterations > 0) logger.log((Index)r.iterations, std::vector<double>{r.final_residual}, 0.0);
        hip::fill_report(report, r);
        logger.summary(report.converged);
        return report.converged;
    }

private:
    Config config_;
    std::shared_ptr<hip::SolverMatrix> mat_;
};

}  // namespace gmres
}  // namespace iterative
}  // namespace solver
}  // namespace linalg
}  // namespace residuum
