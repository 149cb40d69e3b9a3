// rsd_mixed.hpp -- mixed-precision CG state and launchers.
#pragma once

#include "rsd_problem.hpp"

namespace rsd {

// Device state of the mixed solve; only one-block kernels write it.
struct MixedState {
    double res0, res;          // outer FP64 residual norms
    double in_res0, in_tol;    // inner solve: first residual, relative tolerance
    double gam, alp;           // gamma and alpha of the previous inner iteration
    double alpha, beta;        // scalars of the current inner iteration
    int64_t it;                // inner iteration index
    int64_t in_max, total_inner;  // inner iteration cap; SpMVs of all inner solves
    int32_t cycles, converged, stop, asym;  // asym: the FP32 copy's symmetry check failed
    int32_t done;              // the outer loop has ended (converged, asym, or a cap reached)
};

struct MixedWork {
    double* r64 = nullptr;
    float *u = nullptr, *w = nullptr, *p = nullptr, *s = nullptr, *d = nullptr, *invd = nullptr;
    float* ubase = nullptr;   // u - lo: u's buffer with the uniform stencil's padding
    int64_t lo = 0, hi = 0;
    MixedState* st = nullptr;
    // the inner products use the full FP32 copy in the FP64 layout (val32f)
    // instead of the symmetric half storage: for matrices that are not
    // symmetric (a Newton Jacobian with a temperature-dependent conductivity)
    bool full = false;
};

// replays a fixed chunk of inner iterations (a hipGraph in the BE driver)
struct ChunkRunner {
    virtual ~ChunkRunner() = default;
    virtual void run() = 0;
};

// the FP32 copy of A's values; check_symmetry: compare every mirrored pair of
// the half layout (callers whose matrix is symmetric by construction skip it)
void matrix_to_f32(rsd_ctx* ctx, rsd_matrix* A, bool check_symmetry);
// padded: x readable (and finite) from x - uni_lo to x + n + uni_hi, so the
// uniform stencil kernel may run
void launch_spmv_f32(rsd_ctx* ctx, const rsd_matrix* A, const float* x, float* y, bool padded);
// S: the matrix layout the solve uses (its uniform stencil pads u); null
// for none
MixedWork mixed_workspace(rsd_ctx* ctx, int64_t n, const Sell* S);
// the FP32 copy's storage, allocated (and zeroed where the layout keeps
// zeros) once per layout
void alloc_f32(rsd_ctx* ctx, rsd_matrix* A);
// row tiles of 256 per block in the inner vector kernels (set-up and
// update; the SpMV keeps one row per thread): a quarter of the block sums
// for the one-block fold, and more loads in flight per thread
#ifndef RSD_MX_TILES
#define RSD_MX_TILES 4
#endif
constexpr int kMxTiles = RSD_MX_TILES;
int64_t mixed_blocks(int64_t n);
// reduction partials (doubles) a mixed solve on n rows uses
int64_t mixed_partials(int64_t n);
void mixed_launch_inner_iter(rsd_ctx* ctx, const rsd_matrix* A, int32_t precond, const MixedWork& w);
constexpr int kMixedMaxCycles = 40;
constexpr double kMixedDefaultInnerFloor = 1e-4;
// hint: optional [kMixedMaxCycles + 1] inner iteration counts of a previous solve
// (entries 0 .. kMixedMaxCycles are read and written),
// used to queue chunks before the first look at the stop flag (updated)
// ax0 (optional): A x for the x passed in, formed elsewhere bitwise as the
// FP64 row product (the assembly's row pass); the first residual then needs
// no product
MixedState mixed_run(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                     int32_t precond, double inner_floor, bool fresh_f32, const MixedWork& w, ChunkRunner* runner,
                     int64_t* hint = nullptr, const double* ax0 = nullptr);

}  // namespace rsd
