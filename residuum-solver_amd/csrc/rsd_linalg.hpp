// rsd_linalg.hpp -- launchers shared by the solver drivers.
#pragma once

#include "rsd_problem.hpp"

namespace rsd {

// device-resident PCG state (cg/Solver.tpp:15-119 scalars); only the
// one-block fold kernels write it
struct CGState {
    double rz_old, alpha, beta, res0, res;
    int64_t iter;
    int32_t converged, stop;
};

struct CGWork {
    double *r = nullptr, *p = nullptr, *Ap = nullptr, *invd = nullptr, *hist = nullptr;
    CGState* st = nullptr;
    int32_t abs_tol = 0;  // cg::ToleranceType::Absolute: stop on ||r|| < tol instead of the relative test
};

void launch_spmv_f64(rsd_ctx* ctx, const rsd_matrix* A, const double* x, double* y);
void launch_inv_diag(rsd_ctx* ctx, const rsd_matrix* A, double* inv);
// The FP64 stencil copy (Sell::has_st layouts): reserve_stencil allocates
// and zeroes it once per matrix (outside stream capture); once reserved,
// every CG solve on the matrix refreshes it in cg_launch_init and its
// iterations read it.  launch_to_stencil / launch_spmv_stencil need it reserved.
void reserve_stencil(rsd_ctx* ctx, const rsd_matrix* A);
bool stencil_ready(const rsd_matrix* A);
void launch_to_stencil(rsd_ctx* ctx, const rsd_matrix* A);
void launch_spmv_stencil(rsd_ctx* ctx, const rsd_matrix* A, const double* x, double* y);
void launch_dot(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out, bool take_sqrt);
void launch_axpy(rsd_ctx* ctx, int64_t n, double a, const double* x, double* y);
void launch_axpby(rsd_ctx* ctx, int64_t n, double a, double b, const double* x, double* y);
void launch_scal(rsd_ctx* ctx, int64_t n, double a, double* x);
void launch_copy(rsd_ctx* ctx, int64_t n, const double* x, double* y);
void launch_fill(rsd_ctx* ctx, int64_t n, double* p, double v);
void launch_fill_f32(rsd_ctx* ctx, int64_t n, float* p, float v);

CGWork cg_workspace(rsd_ctx* ctx, int64_t n, int64_t max_iters, bool want_hist);
void cg_launch_init(rsd_ctx* ctx, const rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                    int32_t precond, const CGWork& w);
void cg_launch_iter(rsd_ctx* ctx, const rsd_matrix* A, double* x, double tol, int64_t max_iters, int32_t precond,
                    const CGWork& w);
CGState cg_run(rsd_ctx* ctx, const rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
               int32_t precond, const CGWork& w);
// the same iteration around an operator applied by other kernels: init
// with w.Ap = A x (and w.invd set for Jacobi); each iteration with
// w.Ap = A p already formed
void cg_launch_init_ax(rsd_ctx* ctx, int64_t n, const double* b, double tol, int64_t max_iters, int32_t precond,
                       const CGWork& w);
void cg_launch_iter_ap(rsd_ctx* ctx, int64_t n, double* x, double tol, int64_t max_iters, int32_t precond,
                       const CGWork& w);
void cg_report(const CGState& s, int64_t max_iters, rsd_report* rep);

}  // namespace rsd
