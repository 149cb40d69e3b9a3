// rsd_drivers.hip -- solver drivers on the device: the linear Backward
// Euler loop (BackwardEulerRunner::step over BackwardEulerStage).
//
// A step is assemble (K + M/dt, load, M U_prev / dt, essential lift) and a
// warm-started PCG solve.  With use_graph the whole step -- assembly, right
// hand side, CG set-up and the first CG chunk -- is captured once as a
// hipGraph and replayed every step; further CG chunks replay a second graph
// until the device-side stop flag is set.  Capture requires the boundary
// values and the source not to depend on t (their kernels' time argument
// would be frozen into the graph); otherwise the step runs eagerly.
#include "rsd_assembly.hpp"
#include "rsd_linalg.hpp"
#include "rsd_mixed.hpp"
#include "rsd_gmres.hpp"

#include <algorithm>
#include <cmath>

namespace rsd {

namespace {

struct BEBuffers {
    double *Uprev, *Udot, *F, *ms;
};

// BackwardEulerStage::assemble (BackwardEulerStage.hpp:221-246)
// axo (nullable): the row gather pass also forms axo = (K + M/dt) U
void be_assemble(rsd_problem* pb, rsd_matrix* K, const BEBuffers& b, double* U, double ct, double dt,
                 const Expr* src, double* axo = nullptr) {
    rsd_ctx* ctx = pb->ctx;
    const int64_t n = pb->n_free;
    const int32_t kinds[2] = {RSD_FORM_DIFFUSION, RSD_FORM_MASS | RSD_FORM_SCALED};
    const double coefs[2] = {1.0, 1.0 / dt};
    const int32_t kmass[1] = {RSD_FORM_MASS};
    const double one[1] = {1.0};
    launch_copy(ctx, n, U, b.Udot);
    launch_axpby(ctx, n, -1.0 / dt, 1.0 / dt, b.Uprev, b.Udot);
    if (matrix_mass_fusable(pb, 2, kinds)) {
        // K + M/dt and M U_prev from one element pass
        assemble_matrix_mass(pb, ct, 2, kinds, coefs, U, b.Udot, K, b.Uprev, ct - dt, b.ms, axo ? U : nullptr, axo);
        assemble_load(pb, ct, src, nullptr, b.F);
    } else {
        RSD_CHECK(!axo, "be_assemble: the product with the matrix comes only from the fused pass");
        assemble_matrix(pb, ct, 2, kinds, coefs, U, b.Udot, K);
        assemble_load(pb, ct, src, nullptr, b.F);
        assemble_vector(pb, ct - dt, 1, kmass, one, b.Uprev, nullptr, nullptr, RSD_SEL_VALUE, b.ms);
    }
    launch_axpy(ctx, n, 1.0 / dt, b.ms, b.F);
    apply_essential(pb, ct, 2, kinds, coefs, b.F);
}

struct Pinned {
    CGState* p = nullptr;
    Pinned() { RSD_HIP(hipHostMalloc((void**)&p, sizeof(CGState), hipHostMallocDefault)); }
    ~Pinned() { if (p) (void)hipHostFree(p); }
};

}  // namespace

constexpr int64_t kChunk = 8;

// relative tolerance floor of an inner FP32 solve in the mixed BE loop:
// measured at 128^3 over 20 steps, 1e-4 needs the fewest inner SpMVs and the
// least time (1e-3, 3e-4, 1e-5, 3e-6 and 1e-6 are 3-8 % slower)
constexpr double kMixedInnerFloor = kMixedDefaultInnerFloor;

// returns steps taken; throws on a failed solve
struct GraphChunk : ChunkRunner {
    hipStream_t st;
    Graph* g;
    void run() override { RSD_HIP(hipGraphLaunch(g->e, st)); }
};

struct Events {
    hipEvent_t a = nullptr, b = nullptr, c = nullptr;
    bool on = false;
    double asm_ms = 0.0, solve_ms = 0.0;
    explicit Events(bool enable) : on(enable) {
        if (!on) return;
        RSD_HIP(hipEventCreate(&a));
        RSD_HIP(hipEventCreate(&b));
        RSD_HIP(hipEventCreate(&c));
    }
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (c) (void)hipEventDestroy(c);
    }
    void rec(hipEvent_t e, hipStream_t st) { if (on) RSD_HIP(hipEventRecord(e, st)); }
    void accumulate() {
        if (!on) return;
        RSD_HIP(hipEventSynchronize(c));
        float x = 0.f, y = 0.f;
        RSD_HIP(hipEventElapsedTime(&x, a, b));
        RSD_HIP(hipEventElapsedTime(&y, b, c));
        asm_ms += x;
        solve_ms += y;
    }
};

int64_t be_run(rsd_problem* pb, double t0, double tf, double dt_policy, const char* src_text, double tol,
               int64_t max_iters, int32_t precond, int32_t precision, bool use_graph, double* U, int64_t max_steps,
               int64_t* iters, double* timing) {
    Events ev(timing != nullptr);
    rsd_ctx* ctx = pb->ctx;
    const int64_t n = pb->n_free;
    hipStream_t st = ctx->stream;
    const size_t vb = sizeof(double) * (size_t)std::max<int64_t>(n, 1);
    BEBuffers b{(double*)ctx->ws("be.Uprev", vb), (double*)ctx->ws("be.Udot", vb), (double*)ctx->ws("be.F", vb),
                (double*)ctx->ws("be.ms", vb)};
    rsd_matrix K;
    K.pb = pb;
    K.val.alloc(pb->sell.slots);
    const int32_t be_kinds[2] = {RSD_FORM_DIFFUSION, RSD_FORM_MASS | RSD_FORM_SCALED};
    assembly_reserve(pb, 2, be_kinds, true);  // before any capture of the step's assembly
    const Expr* src = cached_source(pb, src_text);
    const bool graphable = use_graph && !pb->bc_time_dependent && !(src && src->uses_t);
    CGWork w = cg_workspace(ctx, n, max_iters, false);
    // every reduction area the captured launches use, reserved before capture
    ctx->partials(std::max(3 * std::max<int64_t>(1, nblocks(n)), mixed_partials(n)));
    const bool mixed = precision == RSD_PREC_MIXED;
    bool sym_model = true;
    if (pb->model.aniso)
        for (int i = 0; i < pb->dim; ++i)
            for (int j = 0; j < i; ++j) sym_model = sym_model && pb->model.tensor[i * pb->dim + j] == pb->model.tensor[j * pb->dim + i];
    MixedWork mw;
    double* ax0 = nullptr;  // (K + M/dt) U from the assembly's row pass: the solve's first product
    if (mixed) {
        build_half(pb);
        mw = mixed_workspace(ctx, n, &pb->sell);
        alloc_f32(ctx, &K);
        const int32_t kinds[2] = {RSD_FORM_DIFFUSION, RSD_FORM_MASS | RSD_FORM_SCALED};
        if (gather_forms_product(pb) && matrix_mass_fusable(pb, 2, kinds)) ax0 = (double*)ctx->ws("be.Ax0", vb);
    }
    // the FP64 CG reads the stencil copy where the layout has one (before capture)
    if (!mixed) reserve_stencil(ctx, &K);
    update_gval(pb, t0);
    launch_copy(ctx, n, U, b.Uprev);
    Pinned pin;
    Graph step_graph, start_graph, chunk_graph;
    double graph_dt = -1.0;
    int64_t last_iters = 0;
    int64_t mixed_hint[kMixedMaxCycles + 1] = {};
    double time = t0;
    int64_t step = 0;
    while (!(time >= tf - 1e-9)) {
        RSD_CHECK(step < max_steps, "be_run: more steps than max_steps");
        const double dt = (dt_policy < tf - time) ? dt_policy : (tf - time);
        const double ct = time + dt;
        if (mixed) {
            ev.rec(ev.a, st);
            if (graphable) {
                if (!step_graph.e || graph_dt != dt) {
                    capture(st, step_graph, [&] { be_assemble(pb, &K, b, U, ct, dt, src, ax0); });
                    graph_dt = dt;
                }
                if (!chunk_graph.e)
                    capture(st, chunk_graph, [&] {
                        for (int64_t k = 0; k < kChunk; ++k) mixed_launch_inner_iter(ctx, &K, precond, mw);
                    });
                RSD_HIP(hipGraphLaunch(step_graph.e, st));
            } else {
                be_assemble(pb, &K, b, U, ct, dt, src, ax0);
            }
            ev.rec(ev.b, st);
            GraphChunk gc;
            gc.st = st;
            gc.g = &chunk_graph;
            // DIFFUSION + MASS/dt with an isotropic or symmetric conductivity is
            // symmetric to rounding: the FP32 conversion skips the pair check
            matrix_to_f32(ctx, &K, !sym_model);
            const MixedState ms = mixed_run(ctx, &K, b.F, U, tol, max_iters, precond, kMixedInnerFloor, true, mw,
                                            graphable ? &gc : nullptr, mixed_hint, ax0);
            ev.rec(ev.c, st);
            ev.accumulate();
            if (iters) iters[step] = ms.total_inner;
            if (!ms.converged)
                throw Error(RSD_ERR_NOT_CONVERGED, "be_run: mixed CG did not converge in step " + std::to_string(step));
            launch_copy(ctx, n, U, b.Uprev);
            time += dt;
            ++step;
            continue;
        }
        ev.rec(ev.a, st);
        if (graphable) {
            if (!step_graph.e || graph_dt != dt) {
                capture(st, step_graph, [&] { be_assemble(pb, &K, b, U, ct, dt, src); });
                capture(st, start_graph, [&] {
                    cg_launch_init(ctx, &K, b.F, U, tol, max_iters, precond, w);
                    for (int64_t k = 0; k < kChunk; ++k) cg_launch_iter(ctx, &K, U, tol, max_iters, precond, w);
                });
                graph_dt = dt;
            }
            if (!chunk_graph.e)
                capture(st, chunk_graph, [&] {
                    for (int64_t k = 0; k < kChunk; ++k) cg_launch_iter(ctx, &K, U, tol, max_iters, precond, w);
                });
            RSD_HIP(hipGraphLaunch(step_graph.e, st));
            ev.rec(ev.b, st);
            RSD_HIP(hipGraphLaunch(start_graph.e, st));
        } else {
            be_assemble(pb, &K, b, U, ct, dt, src);
            ev.rec(ev.b, st);
            cg_launch_init(ctx, &K, b.F, U, tol, max_iters, precond, w);
            for (int64_t k = 0; k < kChunk; ++k) cg_launch_iter(ctx, &K, U, tol, max_iters, precond, w);
        }
        int64_t launched = kChunk;
        auto launch_chunk = [&] {
            if (graphable) RSD_HIP(hipGraphLaunch(chunk_graph.e, st));
            else for (int64_t k = 0; k < kChunk; ++k) cg_launch_iter(ctx, &K, U, tol, max_iters, precond, w);
            launched += kChunk;
        };
        // queue the previous step's iteration count and one more (whole
        // chunks, then single iterations) before the first look at the stop
        // flag: one host round trip per step when the count is steady, and
        // at most a few launches left over after the stop
        if (last_iters > 0) {
            const int64_t want = std::min(last_iters + 1, max_iters);
            while (launched + kChunk <= want) launch_chunk();
            for (; launched < want; ++launched) cg_launch_iter(ctx, &K, U, tol, max_iters, precond, w);
        }
        for (;;) {
            RSD_HIP(hipMemcpyAsync(pin.p, w.st, sizeof(CGState), hipMemcpyDeviceToHost, st));
            RSD_HIP(hipStreamSynchronize(st));
            if (pin.p->stop) break;
            RSD_CHECK(launched <= max_iters + kChunk, "be_run: CG did not stop");
            launch_chunk();
        }
        ev.rec(ev.c, st);
        ev.accumulate();
        const CGState s = *pin.p;
        last_iters = s.iter;
        if (iters) iters[step] = s.converged ? s.iter : max_iters;
        if (!s.converged) throw Error(RSD_ERR_NOT_CONVERGED, "be_run: CG did not converge in step " + std::to_string(step));
        launch_copy(ctx, n, U, b.Uprev);
        time += dt;
        ++step;
    }
    RSD_HIP(hipStreamSynchronize(st));
    if (timing) {
        timing[0] = ev.asm_ms;
        timing[1] = ev.solve_ms;
    }
    return step;
}

// NewtonRunner::solve over SteadyStage (Newton.hpp:32-83, SteadyStage.hpp:60-116)
int64_t newton_run(rsd_problem* pb, const char* src_text, double abs_tol, double rel_tol, int64_t max_newton,
                   int32_t lin_solver, double lin_tol, int64_t lin_max, int32_t krylov, int32_t precond, double* U,
                   double* res_hist, int64_t* lin_iters) {
    rsd_ctx* ctx = pb->ctx;
    const int64_t n = pb->n_free;
    hipStream_t st = ctx->stream;
    const size_t vb = sizeof(double) * (size_t)std::max<int64_t>(n, 1);
    double* F = (double*)ctx->ws("nt.F", vb);
    double* R = (double*)ctx->ws("nt.R", vb);
    double* dU = (double*)ctx->ws("nt.dU", vb);
    double* nrm = (double*)ctx->ws("nt.norm", sizeof(double));
    rsd_matrix J;
    J.pb = pb;
    J.val.alloc(pb->sell.slots);
    const Expr* src = cached_source(pb, src_text);
    const int32_t kdiff[1] = {RSD_FORM_DIFFUSION};
    const int32_t kjac[2] = {RSD_FORM_DIFFUSION, RSD_FORM_TDIFFUSION};
    const double one[2] = {1.0, 1.0};
    double res0 = 0.0;
    for (int64_t it = 0; it < max_newton; ++it) {
        assemble_load(pb, 0.0, src, nullptr, F);
        assemble_vector(pb, 0.0, 1, kdiff, one, U, nullptr, nullptr, RSD_SEL_VALUE, R);
        launch_axpby(ctx, n, 1.0, -1.0, F, R);
        launch_dot(ctx, n, R, R, nrm, true);
        double rn = 0.0;
        RSD_HIP(hipMemcpyAsync(&rn, nrm, sizeof(double), hipMemcpyDeviceToHost, st));
        RSD_HIP(hipStreamSynchronize(st));
        if (res_hist) res_hist[it] = rn;
        if (it == 0) res0 = rn;
        const double rr = res0 > 0.0 ? rn / res0 : 0.0;
        if (rn < abs_tol || rr < rel_tol) return it;
        assemble_matrix(pb, 0.0, 2, kjac, one, U, nullptr, &J);
        launch_fill(ctx, n, dU, 0.0);
        int64_t li = 0;
        bool ok;
        if (lin_solver == 0) {
            CGWork w = cg_workspace(ctx, n, lin_max, false);
            const CGState s = cg_run(ctx, &J, R, dU, lin_tol, lin_max, precond, w);
            ok = s.converged;
            li = s.converged ? s.iter : lin_max;
        } else if (lin_solver == 1) {
            const GMRESResult g = gmres_run(ctx, &J, R, dU, lin_tol, lin_max, krylov, precond);
            ok = g.converged;
            li = g.iterations;
        } else if (lin_solver == 2) {  // FP32 inner GMRES with FP64 refinement
            const MixedGMResult g = gmres_mixed_run(ctx, &J, R, dU, lin_tol, lin_max, krylov, precond,
                                                    kMixedDefaultInnerFloor);
            ok = g.converged;
            li = g.iterations;
        } else {  // FP32 inner CG on the full FP32 copy (J need not be symmetric), FP64 refinement
            MixedWork mw = mixed_workspace(ctx, n, nullptr);
            mw.full = true;
            const MixedState s = mixed_run(ctx, &J, R, dU, lin_tol, lin_max, precond, kMixedDefaultInnerFloor, false,
                                           mw, nullptr);
            ok = s.converged;
            li = s.total_inner;
        }
        if (lin_iters) lin_iters[it] = li;
        if (!ok) return -2;
        launch_axpy(ctx, n, 1.0, dU, U);
    }
    RSD_HIP(hipStreamSynchronize(st));
    return -1;
}

}  // namespace rsd

using namespace rsd;

extern "C" {

int rsd_be_run(rsd_ctx* ctx, rsd_problem* pb, double t0, double tf, double dt, const char* src_expr, double cg_tol,
               int64_t cg_max_iters, int32_t precond, int32_t precision, int32_t use_graph, double* d_U,
               int64_t max_steps, int64_t* steps, int64_t* iters, double* timing_ms) {
    int64_t taken = 0;
    const int rc = guard([&] {
        RSD_CHECK(ctx && pb && d_U, "rsd_be_run: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        RSD_CHECK(dt > 0.0 && tf >= t0, "rsd_be_run: need dt > 0 and tf >= t0");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        RSD_CHECK(precision == RSD_PREC_FP64 || precision == RSD_PREC_MIXED, "unknown precision");
        taken = be_run(pb, t0, tf, dt, src_expr, cg_tol, cg_max_iters, precond, precision, use_graph != 0, d_U,
                       max_steps, iters, timing_ms);
    });
    if (steps) *steps = taken;
    return rc;
}

int rsd_newton_run(rsd_ctx* ctx, rsd_problem* pb, const char* src_expr, double abs_tol, double rel_tol,
                   int64_t max_newton, int32_t lin_solver, double lin_tol, int64_t lin_max_iters, int32_t krylov,
                   int32_t precond, double* d_U, int64_t* newton_iters, double* res_hist, int64_t* lin_iters) {
    int64_t its = -1;
    const int rc = guard([&] {
        RSD_CHECK(ctx && pb && d_U, "rsd_newton_run: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        RSD_CHECK(lin_solver >= 0 && lin_solver <= 3,
                  "lin_solver must be 0 (CG), 1 (GMRES), 2 (mixed GMRES) or 3 (mixed CG)");
        RSD_CHECK(krylov >= 1 && krylov <= 512, "krylov dimension outside [1, 512]");
        its = newton_run(pb, src_expr, abs_tol, rel_tol, max_newton, lin_solver, lin_tol, lin_max_iters, krylov,
                         precond, d_U, res_hist, lin_iters);
    });
    if (newton_iters) *newton_iters = its;
    return rc;
}

}  // extern "C"
