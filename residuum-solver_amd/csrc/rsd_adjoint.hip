// rsd_adjoint.hip -- discrete adjoint of a fixed-dt linear Backward Euler
// run with respect to the per-element log-conductivity (no reference
// counterpart: the reference has no gradients).
//
// Forward:  A u^{k+1} = (1/dt) M u^k|full - lift,  A = K(kappa) + M/dt,
//           every u^k kept in HBM (checkpoints, no recomputation).
// Loss:     L = 1/2 sum_obs (u^N_i - t_i)^2.
// Backward: A lam_N = dL/du^N,  A lam_k = (1/dt) M lam_{k+1}  (A, M symmetric),
//           dL/dtheta_e = -sum_{k=N..1} kappa_e lam_e^k . (Khat_e T_e^k),
//           kappa_e = exp(theta_e), Khat_e the element diffusion matrix
//           with kappa_e taken out (the model's k_scalar and tensor kept),
//           T_e^k the element temperatures with their Dirichlet values.
// The conductivity must not depend on T: the run is linear.  A repeated
// observation index counts once per observation in both L and dL/du^N.
#include "rsd_assembly.hpp"
#include "rsd_linalg.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace rsd {

double adjoint_run(rsd_problem* pb, const double* d_kappa, double t0, double dt, int64_t n_steps, double tol,
                   int64_t max_iters, const double* U0, int64_t n_obs, const int64_t* obs, const double* t_obs,
                   double* grad, double* UN_out, bool want_grad) {
    rsd_ctx* ctx = pb->ctx;
    hipStream_t st = ctx->stream;
    const int64_t n = pb->n_free;
    const size_t vb = sizeof(double) * (size_t)std::max<int64_t>(n, 1);
    const double* saved_ext = pb->kappa_ext;
    pb->kappa_ext = d_kappa;
    struct Restore {
        rsd_problem* pb;
        const double* k;
        ~Restore() { pb->kappa_ext = k; }
    } restore{pb, saved_ext};
    double* Uh = (double*)ctx->ws("adj.Uh", vb * (size_t)(n_steps + 1));
    double* F = (double*)ctx->ws("adj.F", vb);
    double* ms = (double*)ctx->ws("adj.ms", vb);
    double* lam = (double*)ctx->ws("adj.lam", vb);
    double* rhs = (double*)ctx->ws("adj.rhs", vb);
    rsd_matrix A, M;
    A.pb = M.pb = pb;
    A.val.alloc(pb->sell.slots);
    M.val.alloc(pb->sell.slots);
    const int32_t kbe[2] = {RSD_FORM_DIFFUSION, RSD_FORM_MASS | RSD_FORM_SCALED};
    const double cbe[2] = {1.0, 1.0 / dt};
    const int32_t km[1] = {RSD_FORM_MASS};
    const double one[1] = {1.0};
    const double idt = 1.0 / dt;
    assemble_matrix(pb, t0, 2, kbe, cbe, U0, nullptr, &A);
    assemble_matrix(pb, t0, 1, km, one, U0, nullptr, &M);
    launch_copy(ctx, n, U0, Uh);
    CGWork w = cg_workspace(ctx, n, max_iters, false);
    double time = t0;
    for (int64_t k = 0; k < n_steps; ++k) {
        const double ct = time + dt;
        double* Uk = Uh + k * n;
        double* Un = Uh + (k + 1) * n;
        assemble_vector(pb, ct - dt, 1, km, one, Uk, nullptr, nullptr, RSD_SEL_VALUE, ms);
        launch_fill(ctx, n, F, 0.0);
        launch_axpy(ctx, n, idt, ms, F);
        apply_essential(pb, ct, 2, kbe, cbe, F);
        launch_copy(ctx, n, Uk, Un);
        const CGState s = cg_run(ctx, &A, F, Un, tol, max_iters, RSD_PC_JACOBI, w);
        if (!s.converged) throw Error(RSD_ERR_NOT_CONVERGED, "adjoint: forward CG did not converge");
        time = ct;
    }
    const double* UN = Uh + n_steps * n;
    std::vector<double> hU(n);
    RSD_HIP(hipMemcpyAsync(hU.data(), UN, vb, hipMemcpyDeviceToHost, st));
    RSD_HIP(hipStreamSynchronize(st));
    if (UN_out) RSD_HIP(hipMemcpyAsync(UN_out, UN, vb, hipMemcpyDeviceToDevice, st));
    double loss = 0.0;
    std::vector<double> hr(n, 0.0);
    for (int64_t k = 0; k < n_obs; ++k) {
        RSD_CHECK(obs[k] >= 0 && obs[k] < n, "observation index outside the free DOFs");
        const double d = hU[obs[k]] - t_obs[k];
        loss += d * d;
        hr[obs[k]] += d;  // a repeated index contributes once per observation, as in the loss
    }
    loss = 0.5 * loss;
    if (!want_grad) {
        RSD_HIP(hipStreamSynchronize(st));
        return loss;
    }
    RSD_HIP(hipMemcpyAsync(rhs, hr.data(), vb, hipMemcpyHostToDevice, st));
    launch_fill(ctx, pb->n_elem, grad, 0.0);
    for (int64_t k = n_steps; k >= 1; --k) {
        if (k < n_steps) {
            launch_spmv_f64(ctx, &M, lam, F);
            launch_fill(ctx, n, rhs, 0.0);
            launch_axpy(ctx, n, idt, F, rhs);
        }
        launch_fill(ctx, n, lam, 0.0);
        const CGState s = cg_run(ctx, &A, rhs, lam, tol, max_iters, RSD_PC_JACOBI, w);
        if (!s.converged) throw Error(RSD_ERR_NOT_CONVERGED, "adjoint: backward CG did not converge");
        adjoint_grad_step(pb, t0 + (double)k * dt, Uh + k * n, lam, d_kappa, grad);
    }
    RSD_HIP(hipStreamSynchronize(st));
    return loss;
}

}  // namespace rsd

using namespace rsd;

extern "C" int rsd_adjoint_grad_logk(rsd_ctx* ctx, rsd_problem* pb, const double* d_kappa, double t0, double dt,
                                     int64_t n_steps, double cg_tol, int64_t cg_max_iters, const double* d_U0,
                                     int64_t n_obs, const int64_t* obs, const double* t_obs, int32_t want_grad,
                                     double* loss, double* d_grad, double* d_UN) {
    return guard([&] {
        RSD_CHECK(ctx && pb && d_kappa && d_U0 && loss, "rsd_adjoint_grad_logk: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        RSD_CHECK(dt > 0.0 && n_steps >= 1, "need dt > 0 and n_steps >= 1");
        RSD_CHECK(n_obs >= 0 && (n_obs == 0 || (obs && t_obs)), "bad observations");
        RSD_CHECK(!want_grad || d_grad, "gradient output missing");
        RSD_CHECK(!pb->model.has_k && !pb->model.has_dk,
                  "rsd_adjoint_grad_logk: the conductivity must not depend on T (the gradient is for a linear run)");
        *loss = adjoint_run(pb, d_kappa, t0, dt, n_steps, cg_tol, cg_max_iters, d_U0, n_obs, obs, t_obs, d_grad, d_UN,
                            want_grad != 0);
    });
}
