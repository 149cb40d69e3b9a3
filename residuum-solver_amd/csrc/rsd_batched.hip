// rsd_batched.hip -- many independent PCG solves in one set of launches.
//
// The batch concatenates the problems' free-DOF vectors, each segment padded
// to a multiple of 256 rows, so every 256-row block belongs to exactly one
// problem.  A block finds its problem in a block table, walks that problem's
// own sliced-ELL arrays and stores its block partial in that problem's
// partial range; one fold block per problem then finishes the sums.  The
// reduction tree is the one a standalone solve of the problem uses, so each
// batched solve is bitwise identical to rsd_cg_solve_f64 on its own.  Converged problems drop out block by block while the others
// continue; one chunk of iterations for the whole batch is a single graph
// launch.
#include "rsd_batched.hpp"
#include "rsd_linalg.hpp"
#include "rsd_rowdot.hpp"

#include <algorithm>
#include <cmath>

namespace rsd {

// per-problem PCG state (cg/Solver.tpp:15-119 scalars)
struct BState {
    double rz_old, alpha, beta, res0, res;
    int64_t iter;
    int32_t converged, stop;
};

struct BatchDev {
    const int64_t* const* slice_ptr;  // per problem
    const int32_t* const* rp;
    const int32_t* const* col;
    const double* const* val;
    const int32_t* const* diag;
    const int32_t* blk2prob;          // [total blocks]
    const int64_t* first_block;       // [np + 1]
    const int64_t* nrows;             // [np]
    const int64_t* voff;              // [np + 1] vector segment offsets
    __device__ __forceinline__ Cols cols(int pid) const {
        Cols c;
        c.slice_ptr = slice_ptr[pid];
        c.rp = rp[pid];
        c.col = col[pid];
        return c;
    }
};

__device__ __forceinline__ void locate(const BatchDev& B, int64_t gb, int& pid, int64_t& lb, int64_t& nb) {
    pid = B.blk2prob[gb];
    lb = gb - B.first_block[pid];
    nb = B.first_block[pid + 1] - B.first_block[pid];
}

__global__ void k_b_inv_diag(BatchDev B, int64_t total_rows, double* __restrict__ invd) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= total_rows) return;
    int pid; int64_t lb, nb;
    locate(B, g / kBlock, pid, lb, nb);
    const int64_t r = lb * kBlock + (g % kBlock);
    if (r >= B.nrows[pid]) { invd[g] = 1.0; return; }
    const int32_t s = B.diag[pid][r];
    const double d = s >= 0 ? B.val[pid][s] : 0.0;
    invd[g] = (fabs(d) > 1e-300) ? 1.0 / d : 1.0;
}

__global__ void k_b_spmv(BatchDev B, int64_t total_blocks, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t gb = xcd_block(blockIdx.x, total_blocks);
    int pid; int64_t lb, nb;
    locate(B, gb, pid, lb, nb);
    const int64_t r = lb * kBlock + threadIdx.x;
    const int64_t off = B.voff[pid];
    if (r < B.nrows[pid]) y[off + r] = row_dot_f64<COL_ABS>(r, B.cols(pid), B.val[pid], x + off);
}

__global__ __launch_bounds__(kBlock) void k_b_init(BatchDev B, int64_t total_blocks, const double* __restrict__ b,
                                                   const double* __restrict__ Ax, double* __restrict__ r,
                                                   double* __restrict__ p, const double* __restrict__ invd,
                                                   double* partials) {
    const int64_t gb = xcd_block(blockIdx.x, total_blocks);
    int pid; int64_t lb, nb;
    locate(B, gb, pid, lb, nb);
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t i = B.voff[pid] + rl;
    double v[2] = {0.0, 0.0};
    if (rl < B.nrows[pid]) {
        const double ri = b[i] + -1.0 * Ax[i];
        const double zi = invd ? ri * invd[i] : ri;
        r[i] = ri;
        p[i] = zi;
        v[0] = ri * ri;
        v[1] = ri * zi;
    }
    block_partials<2>(v, lb, nb, partials + 2 * B.first_block[pid]);
}

// one block per problem
__global__ __launch_bounds__(kFold) void k_b_init_fold(BatchDev B, const double* __restrict__ partials, BState* st,
                                                        double tol, int64_t max_iters) {
    const int pid = blockIdx.x;
    const int64_t nb = B.first_block[pid + 1] - B.first_block[pid];
    double out[2];
    fold_partials<2>(nb, partials + 2 * B.first_block[pid], out);
    if (threadIdx.x != 0) return;
    BState* s = st + pid;
    const double res0 = sqrt(out[0]);
    s->res0 = res0;
    s->res = res0;
    s->rz_old = out[1];
    s->iter = 0;
    s->converged = 0;
    s->stop = 0;
    if (res0 / (res0 + 1e-50) < tol) { s->converged = 1; s->stop = 1; }
    else if (max_iters <= 1) s->stop = 1;
}

__global__ __launch_bounds__(kBlock) void k_b_spmv_pap(BatchDev B, int64_t total_blocks, const double* __restrict__ p,
                                                       double* __restrict__ Ap, const BState* st, double* partials) {
    const int64_t gb = xcd_block(blockIdx.x, total_blocks);
    int pid; int64_t lb, nb;
    locate(B, gb, pid, lb, nb);
    if (st[pid].stop) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t off = B.voff[pid];
    double v[1] = {0.0};
    if (rl < B.nrows[pid]) {
        const double y = row_dot_f64<COL_ABS>(rl, B.cols(pid), B.val[pid], p + off);
        Ap[off + rl] = y;
        v[0] = p[off + rl] * y;
    }
    block_partials<1>(v, lb, nb, partials + 2 * B.first_block[pid]);
}

__global__ __launch_bounds__(kFold) void k_b_alpha(BatchDev B, const double* __restrict__ partials, BState* st) {
    const int pid = blockIdx.x;
    BState* s = st + pid;
    if (s->stop) return;
    const int64_t nb = B.first_block[pid + 1] - B.first_block[pid];
    double out[1];
    fold_partials<1>(nb, partials + 2 * B.first_block[pid], out);
    if (threadIdx.x == 0) s->alpha = s->rz_old / out[0];
}

__global__ __launch_bounds__(kBlock) void k_b_update_r(BatchDev B, int64_t total_blocks, double* __restrict__ x,
                                                       const double* __restrict__ p, const double* __restrict__ Ap,
                                                       double* __restrict__ r, const double* __restrict__ invd,
                                                       const BState* st, double* partials) {
    const int64_t gb = xcd_block(blockIdx.x, total_blocks);
    int pid; int64_t lb, nb;
    locate(B, gb, pid, lb, nb);
    const BState* s = st + pid;
    if (s->stop) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t i = B.voff[pid] + rl;
    const double alpha = s->alpha;
    double v[2] = {0.0, 0.0};
    if (rl < B.nrows[pid]) {
        x[i] += alpha * p[i];
        const double ri = r[i] + -alpha * Ap[i];
        r[i] = ri;
        const double zi = invd ? ri * invd[i] : ri;
        v[0] = ri * ri;
        v[1] = ri * zi;
    }
    block_partials<2>(v, lb, nb, partials + 2 * B.first_block[pid]);
}

__global__ __launch_bounds__(kFold) void k_b_beta(BatchDev B, const double* __restrict__ partials, BState* st,
                                                   double tol, int64_t max_iters) {
    const int pid = blockIdx.x;
    BState* s = st + pid;
    if (s->stop) return;
    const int64_t nb = B.first_block[pid + 1] - B.first_block[pid];
    double out[2];
    fold_partials<2>(nb, partials + 2 * B.first_block[pid], out);
    if (threadIdx.x != 0) return;
    const int64_t k = s->iter + 1;
    s->iter = k;
    const double res = sqrt(out[0]);
    s->res = res;
    if (res / (s->res0 + 1e-50) < tol) { s->converged = 1; s->stop = 1; }
    else {
        s->beta = out[1] / s->rz_old;
        s->rz_old = out[1];
        if (k >= max_iters - 1) s->stop = 1;
    }
}

__global__ __launch_bounds__(kBlock) void k_b_update_p(BatchDev B, int64_t total_blocks, double* __restrict__ p,
                                                       const double* __restrict__ r, const double* __restrict__ invd,
                                                       const BState* st) {
    const int64_t gb = xcd_block(blockIdx.x, total_blocks);
    int pid; int64_t lb, nb;
    locate(B, gb, pid, lb, nb);
    const BState* s = st + pid;
    if (s->stop) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t i = B.voff[pid] + rl;
    const double beta = s->beta;
    if (rl < B.nrows[pid]) {
        const double zi = invd ? r[i] * invd[i] : r[i];
        p[i] = beta * p[i] + 1.0 * zi;
    }
}

}  // namespace rsd

using namespace rsd;

struct rsd_batch {
    rsd_ctx* ctx = nullptr;
    int32_t np = 0;
    int64_t total_rows = 0, total_blocks = 0;
    std::vector<int64_t> h_voff, h_nrows, h_first;
    DBuf<const int64_t*> d_slice;
    DBuf<const int32_t*> d_rp, d_col, d_diag;
    DBuf<const double*> d_val;
    DBuf<int32_t> blk2prob;
    DBuf<int64_t> first_block, nrows, voff;
    DBuf<double> r, p, Ap, invd;
    DBuf<BState> st;
    DBuf<double> partials;
    BatchDev dev() const {
        BatchDev B;
        B.slice_ptr = d_slice.p;
        B.rp = d_rp.p;
        B.col = d_col.p;
        B.val = d_val.p;
        B.diag = d_diag.p;
        B.blk2prob = blk2prob.p;
        B.first_block = first_block.p;
        B.nrows = nrows.p;
        B.voff = voff.p;
        return B;
    }
};

namespace rsd {

static void b_launch_iter(rsd_batch* bt, double* x, double tol, int64_t max_iters, int32_t precond) {
    hipStream_t st = bt->ctx->stream;
    const BatchDev B = bt->dev();
    const int64_t nb = bt->total_blocks;
    const double* invd = precond == RSD_PC_JACOBI ? bt->invd.p : nullptr;
    const unsigned np = (unsigned)bt->np;
    hipLaunchKernelGGL(k_b_spmv_pap, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, bt->p.p, bt->Ap.p, bt->st.p,
                       bt->partials.p);
    hipLaunchKernelGGL(k_b_alpha, dim3(np), dim3(kFold), 0, st, B, bt->partials.p, bt->st.p);
    hipLaunchKernelGGL(k_b_update_r, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, x, bt->p.p, bt->Ap.p, bt->r.p,
                       invd, bt->st.p, bt->partials.p);
    hipLaunchKernelGGL(k_b_beta, dim3(np), dim3(kFold), 0, st, B, bt->partials.p, bt->st.p, tol, max_iters);
    hipLaunchKernelGGL(k_b_update_p, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, bt->p.p, bt->r.p, invd, bt->st.p);
}

void batched_cg(rsd_batch* bt, const double* b, double* x, double tol, int64_t max_iters, int32_t precond,
                bool use_graph, std::vector<BState>& out, int64_t chunk) {
    hipStream_t st = bt->ctx->stream;
    const BatchDev B = bt->dev();
    const int64_t nb = bt->total_blocks;
    if (precond == RSD_PC_JACOBI)
        hipLaunchKernelGGL(k_b_inv_diag, dim3((unsigned)nb), dim3(kBlock), 0, st, B, bt->total_rows, bt->invd.p);
    hipLaunchKernelGGL(k_b_spmv, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, x, bt->Ap.p);
    hipLaunchKernelGGL(k_b_init, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, b, bt->Ap.p, bt->r.p, bt->p.p,
                       precond == RSD_PC_JACOBI ? bt->invd.p : nullptr, bt->partials.p);
    hipLaunchKernelGGL(k_b_init_fold, dim3((unsigned)bt->np), dim3(kFold), 0, st, B, bt->partials.p, bt->st.p, tol,
                       max_iters);
    RSD_HIP(hipGetLastError());
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (use_graph) {
        RSD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int64_t k = 0; k < chunk; ++k) b_launch_iter(bt, x, tol, max_iters, precond);
        RSD_HIP(hipStreamEndCapture(st, &g));
        RSD_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    }
    BState* h = nullptr;
    RSD_HIP(hipHostMalloc((void**)&h, sizeof(BState) * (size_t)bt->np, hipHostMallocDefault));
    int64_t launched = 0;
    try {
        for (;;) {
            RSD_HIP(hipMemcpyAsync(h, bt->st.p, sizeof(BState) * (size_t)bt->np, hipMemcpyDeviceToHost, st));
            RSD_HIP(hipStreamSynchronize(st));
            bool all = true;
            for (int i = 0; i < bt->np; ++i) all = all && h[i].stop;
            if (all) break;
            RSD_CHECK(launched <= max_iters + chunk, "batched CG did not stop");
            if (ge) RSD_HIP(hipGraphLaunch(ge, st));
            else for (int64_t k = 0; k < chunk; ++k) b_launch_iter(bt, x, tol, max_iters, precond);
            RSD_HIP(hipGetLastError());
            launched += chunk;
        }
        out.assign(h, h + bt->np);
    } catch (...) {
        (void)hipHostFree(h);
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    (void)hipHostFree(h);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
}

}  // namespace rsd

extern "C" {

int rsd_batch_create(rsd_ctx* ctx, int32_t n_problems, rsd_matrix* const* mats, rsd_batch** out) {
    return guard([&] {
        RSD_CHECK(ctx && out && n_problems > 0 && mats, "rsd_batch_create: bad argument");
        auto bt = std::make_unique<rsd_batch>();
        bt->ctx = ctx;
        bt->np = n_problems;
        std::vector<const int64_t*> sl(n_problems);
        std::vector<const int32_t*> rp(n_problems), col(n_problems), dg(n_problems);
        std::vector<const double*> val(n_problems);
        bt->h_voff.assign(n_problems + 1, 0);
        bt->h_first.assign(n_problems + 1, 0);
        bt->h_nrows.assign(n_problems, 0);
        for (int i = 0; i < n_problems; ++i) {
            RSD_CHECK(mats[i] && mats[i]->pb->ctx == ctx, "rsd_batch_create: matrix on another context");
            const Sell& S = mats[i]->pb->sell;
            sl[i] = S.slice_ptr.p;
            rp[i] = S.row_ptr.p;
            col[i] = S.col.p;
            dg[i] = S.diag.p;
            val[i] = mats[i]->val.p;
            const int64_t nbk = std::max<int64_t>(1, nblocks(S.n_rows));
            bt->h_nrows[i] = S.n_rows;
            bt->h_first[i + 1] = bt->h_first[i] + nbk;
            bt->h_voff[i + 1] = bt->h_voff[i] + nbk * kBlock;
        }
        bt->total_blocks = bt->h_first[n_problems];
        bt->total_rows = bt->h_voff[n_problems];
        std::vector<int32_t> b2p(bt->total_blocks);
        for (int i = 0; i < n_problems; ++i)
            for (int64_t b = bt->h_first[i]; b < bt->h_first[i + 1]; ++b) b2p[b] = i;
        hipStream_t st = ctx->stream;
        bt->d_slice.upload(sl.data(), n_problems, st);
        bt->d_rp.upload(rp.data(), n_problems, st);
        bt->d_col.upload(col.data(), n_problems, st);
        bt->d_diag.upload(dg.data(), n_problems, st);
        bt->d_val.upload(val.data(), n_problems, st);
        bt->blk2prob.upload(b2p.data(), bt->total_blocks, st);
        bt->first_block.upload(bt->h_first.data(), n_problems + 1, st);
        bt->nrows.upload(bt->h_nrows.data(), n_problems, st);
        bt->voff.upload(bt->h_voff.data(), n_problems + 1, st);
        bt->r.alloc(bt->total_rows);
        bt->p.alloc(bt->total_rows);
        bt->Ap.alloc(bt->total_rows);
        bt->invd.alloc(bt->total_rows);
        bt->st.alloc(n_problems);
        bt->partials.alloc(2 * bt->total_blocks);
        RSD_HIP(hipMemsetAsync(bt->r.p, 0, sizeof(double) * bt->total_rows, st));
        RSD_HIP(hipMemsetAsync(bt->p.p, 0, sizeof(double) * bt->total_rows, st));
        RSD_HIP(hipMemsetAsync(bt->Ap.p, 0, sizeof(double) * bt->total_rows, st));
        RSD_HIP(hipStreamSynchronize(st));
        *out = bt.release();
    });
}

int rsd_batch_destroy(rsd_batch* bt) {
    return guard([&] {
        if (!bt) return;
        (void)hipStreamSynchronize(bt->ctx->stream);
        delete bt;
    });
}

int rsd_batch_layout(rsd_batch* bt, int64_t* total_rows, int64_t* offsets) {
    return guard([&] {
        RSD_CHECK(bt, "null batch");
        if (total_rows) *total_rows = bt->total_rows;
        if (offsets) std::copy(bt->h_voff.begin(), bt->h_voff.end(), offsets);
    });
}

int rsd_cg_solve_batched(rsd_ctx* ctx, rsd_batch* bt, const double* b, double* x, double tol, int64_t max_iters,
                         int32_t precond, int32_t use_graph, rsd_report* reps) {
    return guard([&] {
        RSD_CHECK(ctx && bt && b && x && ctx == bt->ctx, "rsd_cg_solve_batched: bad argument");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        std::vector<BState> s;
        batched_cg(bt, b, x, tol, max_iters, precond, use_graph != 0, s, 8);
        if (reps)
            for (int i = 0; i < bt->np; ++i) {
                CGState c{};
                c.res0 = s[i].res0;
                c.res = s[i].res;
                c.iter = s[i].iter;
                c.converged = s[i].converged;
                c.stop = s[i].stop;
                cg_report(c, max_iters, &reps[i]);
            }
    });
}

}  // extern "C"
