// rsd_gmres.hpp -- device GMRES(m) state and driver.
#pragma once

#include "rsd_problem.hpp"

namespace rsd {

struct GMState {
    double beta, res0, hnext, est;
    int64_t total;
    int32_t stop, breakdown, converged, mloc;
};

struct GMRESResult {
    int32_t converged;
    int64_t iterations;
    double initial_residual, final_residual, final_residual_rel;
};

GMRESResult gmres_run(rsd_ctx* ctx, const rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                      int32_t m, int32_t precond, bool abs_tol = false);

// mixed precision: FP64 outer refinement, FP32 inner GMRES(m) on the full
// FP32 copy; iterations counts inner Arnoldi steps, cycles outer refinements
struct MixedGMResult {
    int32_t converged;
    int64_t iterations, cycles;
    double initial_residual, final_residual, final_residual_rel;
};
MixedGMResult gmres_mixed_run(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol,
                              int64_t max_iters, int32_t m, int32_t precond, double inner_floor);

}  // namespace rsd
