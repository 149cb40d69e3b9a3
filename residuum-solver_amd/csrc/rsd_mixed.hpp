// rsd_mixed.hpp -- mixed-precision CG state and launchers.
#pragma once

#include "rsd_problem.hpp"

namespace rsd {

struct MixedState {
    double res0, res;          // outer FP64 residual norms
    double in_res0, in_tol;    // inner solve
    double rz_old, alpha, beta;
    int64_t in_iter, in_max, total_inner;
    int32_t cycles, converged, stop, pad;
};

struct MixedWork {
    double* r64 = nullptr;
    float *r = nullptr, *p = nullptr, *Ap = nullptr, *d = nullptr, *invd = nullptr;
    MixedState* st = nullptr;
};

// replays a fixed chunk of inner iterations (a hipGraph in the BE driver)
struct ChunkRunner {
    virtual ~ChunkRunner() = default;
    virtual void run() = 0;
};

void matrix_to_f32(rsd_ctx* ctx, rsd_matrix* A);
void launch_spmv_f32(rsd_ctx* ctx, const rsd_matrix* A, const float* x, float* y);
MixedWork mixed_workspace(rsd_ctx* ctx, int64_t n);
void mixed_launch_inner_iter(rsd_ctx* ctx, const rsd_matrix* A, int32_t precond, const MixedWork& w);
MixedState mixed_run(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                     int32_t precond, double inner_floor, bool fresh_f32, const MixedWork& w, ChunkRunner* runner);

}  // namespace rsd
