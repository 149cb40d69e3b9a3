// rsd_fem.hip -- CG over the matrix-free operator (FEMOperator).
//
// linalg/operator/FEMOperator.hpp:27-71 applies the operator by element
// assembly: apply(x) = assembleVector with the forms acting on the operand x
// (free-DOF gather), diagonal() = assembleDiagonal.  Here the apply is the
// colour-ordered vector kernels with the operand selector, and the solver
// is the reference CG (cg/Solver.tpp:15-119) with the same device state and
// stop flag as the CSR solver: the vector kernels of an apply return at once
// after the stop flag is set, so chunks of iterations can be queued ahead.
#include "rsd_assembly.hpp"
#include "rsd_linalg.hpp"

#include <algorithm>

namespace rsd {

__global__ void k_jacobi_rule(int64_t n, double* __restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double v = d[i];
    d[i] = (fabs(v) > 1e-300) ? 1.0 / v : 1.0;  // Jacobi.hpp:33-36
}

struct FemOp {
    rsd_problem* pb;
    double time;
    int32_t nf;
    const int32_t* kinds;
    const double* coefs;
    const double* U;
    const double* Udot;

    // FEMOperator::apply (FEMOperator.hpp:38-43)
    void apply(const double* x, double* y, const int32_t* stop) const {
        assemble_vector(pb, time, nf, kinds, coefs, U, Udot, x, RSD_SEL_OPERAND, y, stop);
    }
};

static CGState fem_cg_run(rsd_ctx* ctx, const FemOp& op, const double* b, double* x, double tol, int64_t max_iters,
                          int32_t precond, const CGWork& w) {
    const int64_t n = op.pb->n_free;
    if (precond == RSD_PC_JACOBI) {
        // FEMOperator::diagonal + Jacobi::update
        assemble_diagonal(op.pb, op.time, op.nf, op.kinds, op.coefs, op.U, w.invd);
        if (n) hipLaunchKernelGGL(k_jacobi_rule, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, ctx->stream, n, w.invd);
    }
    op.apply(x, w.Ap, nullptr);
    cg_launch_init_ax(ctx, n, b, tol, max_iters, precond, w);
    CGState h{};
    CGState* pinned = nullptr;
    RSD_HIP(hipHostMalloc((void**)&pinned, sizeof(CGState), hipHostMallocDefault));
    int64_t chunk = 2, done_iters = 0;
    try {
        for (;;) {
            RSD_HIP(hipMemcpyAsync(pinned, w.st, sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
            RSD_HIP(hipStreamSynchronize(ctx->stream));
            h = *pinned;
            if (h.stop) break;
            RSD_CHECK(done_iters <= max_iters + 2, "CG did not stop within max_iters launches");
            const int64_t c = std::min<int64_t>(chunk, std::max<int64_t>(1, max_iters - done_iters));
            for (int64_t k = 0; k < c; ++k) {
                op.apply(w.p, w.Ap, &w.st->stop);
                cg_launch_iter_ap(ctx, n, x, tol, max_iters, precond, w);
            }
            RSD_HIP(hipGetLastError());
            done_iters += c;
            chunk = std::min<int64_t>(chunk * 2, 16);
        }
    } catch (...) {
        (void)hipHostFree(pinned);
        throw;
    }
    (void)hipHostFree(pinned);
    return h;
}

}  // namespace rsd

using namespace rsd;

extern "C" {

int rsd_cg_solve_fem(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t nf, const int32_t* kinds,
                     const double* coefs, const double* U, const double* Udot, const double* b, double* x, double tol,
                     int32_t tol_type, int64_t max_iters, int32_t precond, rsd_report* rep, double* hist) {
    return guard([&] {
        RSD_CHECK(ctx && pb && b && x && (nf == 0 || kinds), "rsd_cg_solve_fem: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        RSD_CHECK(tol_type == RSD_TOL_RELATIVE || tol_type == RSD_TOL_ABSOLUTE, "unknown tolerance type");
        RSD_CHECK(max_iters >= 0, "max_iters must be >= 0");
        const int64_t n = pb->n_free;
        CGWork w = cg_workspace(ctx, n, max_iters, hist != nullptr);
        w.abs_tol = tol_type == RSD_TOL_ABSOLUTE;
        const FemOp op{pb, time, nf, kinds, coefs, U, Udot};
        CGState s = fem_cg_run(ctx, op, b, x, tol, max_iters, precond, w);
        cg_report(s, max_iters, rep);
        if (hist) {
            const int64_t cnt = std::min<int64_t>(s.iter + 1, max_iters + 1);
            RSD_HIP(hipMemcpyAsync(hist, w.hist, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost, ctx->stream));
            RSD_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

}  // extern "C"
