// rsd_batched.hip -- many independent PCG solves in one set of launches.
//
// The batch concatenates the problems' free-DOF vectors, each segment padded
// to a multiple of 256 rows, so every 256-row block belongs to exactly one
// problem.  A block finds its problem in a block table, walks that problem's
// own sliced-ELL arrays and stores its block partial in that problem's
// partial range; one fold block per problem then finishes the sums.  The
// reduction tree is the one a standalone solve of the problem uses, so each
// batched solve is bitwise identical to rsd_cg_solve_f64 on its own.
// Converged problems drop out block by block while the others continue; one
// chunk of iterations for the whole batch is a single graph launch.
//
// Schedule: the chunk length, the row product's load batch, the order in
// which the launch grid visits the problems' blocks (problem by problem, or
// interleaved one block per problem at a time) and the XCD block remap.
// None of them changes an operation or its order, so every schedule gives
// the same bits; rsd_batch_autotune times them on the batch and keeps the
// fastest, cached in a JSON file keyed by the device and the batch.
#include "rsd_batched.hpp"
#include "rsd_linalg.hpp"
#include "rsd_mixed.hpp"
#include "rsd_rowdot.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

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
    const uint8_t* const* pid;        // per problem, shared offset patterns (COL_PAT)
    const int32_t* const* poff;
    const int32_t* const* plen;
    const int32_t* stride;            // [np]
    const double* const* val;
    const int32_t* const* diag;
    const int32_t* blk2prob;          // [total blocks] problem of each launched block
    const int32_t* blk2local;         // [total blocks] its block index within the problem
    const int64_t* first_block;       // [np + 1]
    const int64_t* nrows;             // [np]
    const int64_t* voff;              // [np + 1] vector segment offsets
    const double* const* val_st;      // per problem, its FP64 stencil copy (when every problem has one)
    const uint32_t* const* st_mask;
    const int32_t* st_coff;           // [np * kStRuns] run centre offsets
    int32_t xcd;                      // remap launched blocks XCD by XCD
    __device__ __forceinline__ Cols cols(int p) const {
        Cols c;
        c.slice_ptr = slice_ptr[p];
        c.rp = rp[p];
        c.col = col[p];
        c.pid = pid[p];
        c.poff = poff[p];
        c.plen = plen[p];
        c.stride = stride[p];
        return c;
    }
};

__device__ __forceinline__ void locate(const BatchDev& B, int64_t total_blocks, int& pid, int64_t& lb, int64_t& nb) {
    const int64_t gb = B.xcd ? xcd_block(blockIdx.x, total_blocks) : (int64_t)blockIdx.x;
    pid = B.blk2prob[gb];
    lb = B.blk2local[gb];
    nb = B.first_block[pid + 1] - B.first_block[pid];
}

__global__ void k_b_inv_diag(BatchDev B, int64_t total_blocks, double* __restrict__ invd) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    const int64_t r = lb * kBlock + threadIdx.x;
    const int64_t g = B.voff[pid] + r;
    if (r >= B.nrows[pid]) { invd[g] = 1.0; return; }
    const int32_t s = B.diag[pid][r];
    const double d = s >= 0 ? B.val[pid][s] : 0.0;
    invd[g] = (fabs(d) > 1e-300) ? 1.0 / d : 1.0;
}

template<int MODE, int U>
__global__ void k_b_spmv(BatchDev B, int64_t total_blocks, const double* __restrict__ x, double* __restrict__ y) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    const int64_t r = lb * kBlock + threadIdx.x;
    const int64_t off = B.voff[pid];
    if (r < B.nrows[pid]) y[off + r] = row_dot_f64<MODE, U>(r, B.cols(pid), B.val[pid], x + off);
}

__global__ __launch_bounds__(kBlock) void k_b_init(BatchDev B, int64_t total_blocks, const double* __restrict__ b,
                                                   const double* __restrict__ Ax, double* __restrict__ r,
                                                   double* __restrict__ p, const double* __restrict__ invd,
                                                   double* partials) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
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

constexpr int kStRuns = kStMaxPos / 3;

// k_b_spmv_pap over the problems' stencil copies (k_cg_spmv_pap_st per
// problem; every lane of a block runs the row product)
template<int NR>
__global__ __launch_bounds__(kBlock) void k_b_spmv_pap_st(BatchDev B, int64_t total_blocks,
                                                          const double* __restrict__ p, double* __restrict__ Ap,
                                                          const BState* st, double* partials) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    if (st[pid].stop) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t off = B.voff[pid], n = B.nrows[pid];
    StCols c;
    c.mask = B.st_mask[pid];
    c.np = 3 * NR;
#pragma unroll
    for (int k = 0; k < NR; ++k) c.coff[k] = B.st_coff[pid * kStRuns + k];
    const double y = row_dot_st<NR>(rl, n, c, B.val_st[pid], p + off);
    double v[1] = {0.0};
    if (rl < n) {
        Ap[off + rl] = y;
        v[0] = p[off + rl] * y;
    }
    block_partials<1>(v, lb, nb, partials + 2 * B.first_block[pid]);
}

template<int MODE, int U>
__global__ __launch_bounds__(kBlock) void k_b_spmv_pap(BatchDev B, int64_t total_blocks, const double* __restrict__ p,
                                                       double* __restrict__ Ap, const BState* st, double* partials) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    if (st[pid].stop) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t off = B.voff[pid];
    double v[1] = {0.0};
    if (rl < B.nrows[pid]) {
        const double y = row_dot_f64<MODE, U>(rl, B.cols(pid), B.val[pid], p + off);
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
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
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
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
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

// ------------------------------------------------------------------------
// Batched mixed precision: each problem runs the standalone mixed solve
// (rsd_mixed.hip) -- FP64 outer residual, FP32 inner CG on the problem's
// FP32 copy -- with the outer loop's decisions (first residual,
// convergence, inner tolerance and cap) made per problem on the device, in
// the host's double arithmetic.  The FP64 kernels use the 256-row block
// table, the inner set-up and update the 1024-row tile table of a
// standalone mixed solve; each problem's partials sit in its own range,
// laid out as a standalone solve's, so every problem is bitwise its
// rsd_cg_solve_mixed.  The FP32 products walk the half layout's per-pattern
// tables, bitwise the uniform stencil kernel a standalone solve uses.
// ------------------------------------------------------------------------
struct BMState {
    MixedState m;
    int32_t done;  // the problem's outer loop has ended (converged or given up)
};

struct BMixDev {
    const Cols* c32;             // [np] FP32 columns
    const float* const* val32;   // [np]
    const int32_t* t2prob;       // [total tile blocks]
    const int32_t* t2local;
    const int64_t* first_t;      // [np + 1]
    const int64_t* pbase;        // [np] partial range: nb block sums, then 2 nbu
    const int64_t* off32;        // [np] FP32 vector segment starts (zeros around each, see bm_setup)
};

__device__ __forceinline__ void locate_t(const BatchDev& B, const BMixDev& M, int64_t total_t, int& pid, int64_t& lb,
                                         int64_t& nbu) {
    const int64_t gb = B.xcd ? xcd_block(blockIdx.x, total_t) : (int64_t)blockIdx.x;
    pid = M.t2prob[gb];
    lb = M.t2local[gb];
    nbu = M.first_t[pid + 1] - M.first_t[pid];
}

__global__ void k_bm_inv_diag(BatchDev B, BMixDev M, int64_t total_blocks, float* __restrict__ invd) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    const int64_t r = lb * kBlock + threadIdx.x;
    const int64_t g = M.off32[pid] + r;
    if (r >= B.nrows[pid]) { invd[g] = 1.0f; return; }
    const int32_t s = B.diag[pid][r];
    const double d = s >= 0 ? B.val[pid][s] : 0.0;
    invd[g] = (float)((fabs(d) > 1e-300) ? 1.0 / d : 1.0);
}

// r = b - A x (FP64), block sums of ||r||^2
template<int MODE, int U>
__global__ __launch_bounds__(kBlock) void k_bm_residual(BatchDev B, BMixDev M, int64_t total_blocks,
                                                        const double* __restrict__ b, const double* __restrict__ x,
                                                        double* __restrict__ r64, const BMState* st, double* part) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    if (st[pid].done) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t off = B.voff[pid];
    double v[1] = {0.0};
    if (rl < B.nrows[pid]) {
        const double ri = b[off + rl] + -1.0 * row_dot_f64<MODE, U>(rl, B.cols(pid), B.val[pid], x + off);
        r64[off + rl] = ri;
        v[0] = ri * ri;
    }
    block_partials<1>(v, lb, nb, part + M.pbase[pid]);
}

// one block per problem: the outer loop's step of the standalone solve
__global__ __launch_bounds__(kFold) void k_bm_outer(BatchDev B, BMixDev M, const double* __restrict__ part,
                                                    BMState* st, double tol, int64_t max_iters, double inner_floor) {
    const int pid = blockIdx.x;
    BMState* s = st + pid;
    if (s->done) return;
    const int64_t nb = B.first_block[pid + 1] - B.first_block[pid];
    double out[1];
    fold_partials<1>(nb, part + M.pbase[pid], out);
    if (threadIdx.x != 0) return;
    MixedState& m = s->m;
    const double res = sqrt(out[0]);
    if (m.cycles == 0) m.res0 = res;
    m.res = res;
    const double rel = m.res / (m.res0 + 1e-50);
    if (rel < tol) {
        m.converged = 1;
        m.stop = 1;
        s->done = 1;
        return;
    }
    if (m.cycles >= kMixedMaxCycles || m.total_inner >= max_iters || !(rel == rel)) {
        m.stop = 1;
        s->done = 1;
        return;
    }
    double it = 0.5 * tol * (m.res0 + 1e-50) / m.res;
    it = (inner_floor < it) ? it : inner_floor;  // max(inner_floor, it)
    it = (it < 0.5) ? it : 0.5;                  // min(0.5, it)
    m.in_tol = it;
    m.in_max = max_iters - m.total_inner > 1 ? max_iters - m.total_inner : 1;
    m.cycles = m.cycles + 1;
    m.it = 0;
    m.stop = 0;
}

// inner set-up (k_mx_init per problem)
__global__ __launch_bounds__(kBlock) void k_bm_init(BatchDev B, BMixDev M, int64_t total_t,
                                                    const double* __restrict__ r64, float* __restrict__ u,
                                                    float* __restrict__ p, float* __restrict__ s, float* __restrict__ d,
                                                    const float* __restrict__ invd, const BMState* st, double* part) {
    int pid; int64_t lb, nbu;
    locate_t(B, M, total_t, pid, lb, nbu);
    const BMState* bs = st + pid;
    if (bs->done) return;
    const int64_t n = B.nrows[pid], off = B.voff[pid], off32 = M.off32[pid];
    const double sc = bs->m.res > 0.0 ? 1.0 / bs->m.res : 0.0;
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int t = 0; t < kMxTiles; ++t) {
        const int64_t rl = (lb * kMxTiles + t) * kBlock + threadIdx.x;
        if (rl < n) {
            const int64_t i = off32 + rl;
            const float ri = (float)(r64[off + rl] * sc);
            const float ui = invd ? ri * invd[i] : ri;
            u[i] = ui;
            p[i] = 0.0f;
            s[i] = 0.0f;
            d[i] = 0.0f;
            v[0] += (double)ri * (double)ui;
            v[1] += (double)ri * (double)ri;
        }
    }
    const int64_t nb = B.first_block[pid + 1] - B.first_block[pid];
    block_partials<2>(v, lb, nbu, part + M.pbase[pid] + nb);
}

// w = A u (FP32), block sums of w.u (k_mx_spmv_wu per problem)
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_bm_spmv_wu(BatchDev B, BMixDev M, int64_t total_blocks,
                                                       const float* __restrict__ u, float* __restrict__ w,
                                                       const BMState* st, double* part) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    if (st[pid].m.stop) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    const int64_t off = M.off32[pid];
    double v[1] = {0.0};
    if (rl < B.nrows[pid]) {
        const float wi = row_dot_f32<MODE>(rl, M.c32[pid], M.val32[pid], u + off);
        w[off + rl] = wi;
        v[0] = (double)wi * (double)u[off + rl];
    }
    block_partials<1>(v, lb, nb, part + M.pbase[pid]);
}

// one block per problem: k_mx_scalars
__global__ __launch_bounds__(kFold) void k_bm_scalars(BatchDev B, BMixDev M, const double* __restrict__ part,
                                                      BMState* st) {
    const int pid = blockIdx.x;
    MixedState* ms = &st[pid].m;
    if (ms->stop) return;
    const int64_t nb = B.first_block[pid + 1] - B.first_block[pid], nbu = M.first_t[pid + 1] - M.first_t[pid];
    const double* P0 = part + M.pbase[pid];
    const double* P12 = P0 + nb;
    const int64_t k = ms->it;
    const double gam_prev = ms->gam, alp_prev = ms->alp, in_res0 = ms->in_res0, in_tol = ms->in_tol;
    const int64_t in_max = ms->in_max;
    double o[3];
    const double* pk[3] = {P0, P12, P12 + nbu};
    const int64_t nbk[3] = {nb, nbu, nbu};
    fold_partials<3>(pk, nbk, o);
    if (threadIdx.x != 0) return;
    const double del = o[0], gam = o[1], res = sqrt(o[2]);
    const double r0 = k == 0 ? res : in_res0;
    ms->total_inner += 1;
    if (k == 0) ms->in_res0 = res;
    if (res / (r0 + 1e-50) < in_tol || !(res == res) || k >= in_max) {
        ms->stop = 1;
        return;
    }
    double alpha, beta;
    if (k == 0) {
        beta = 0.0;
        alpha = gam / del;
    } else {
        beta = gam / gam_prev;
        alpha = gam / (del - beta * gam / alp_prev);
    }
    ms->alpha = alpha;
    ms->beta = beta;
    ms->gam = gam;
    ms->alp = alpha;
    ms->it = k + 1;
}

// k_mx_update per problem
__global__ __launch_bounds__(kBlock) void k_bm_update(BatchDev B, BMixDev M, int64_t total_t, float* __restrict__ p,
                                                      float* __restrict__ s, float* __restrict__ d,
                                                      float* __restrict__ u, const float* __restrict__ w,
                                                      const float* __restrict__ invd, const BMState* st,
                                                      double* part) {
    int pid; int64_t lb, nbu;
    locate_t(B, M, total_t, pid, lb, nbu);
    const MixedState* ms = &st[pid].m;
    if (ms->stop) return;
    const int64_t n = B.nrows[pid], off = M.off32[pid];
    const float a = (float)ms->alpha, bt = (float)ms->beta;
    constexpr int T = kMxTiles;
    int64_t idx[T];
    float pv[T], sv[T], dv[T], uv[T], wv[T], mv[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int64_t rl = (lb * T + t) * kBlock + threadIdx.x;
        const bool in = rl < n;
        idx[t] = in ? off + rl : -1;
        pv[t] = in ? p[off + rl] : 0.0f;
        sv[t] = in ? s[off + rl] : 0.0f;
        dv[t] = in ? d[off + rl] : 0.0f;
        uv[t] = in ? u[off + rl] : 0.0f;
        wv[t] = in ? w[off + rl] : 0.0f;
        mv[t] = (in && invd) ? invd[off + rl] : 1.0f;
    }
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int t = 0; t < T; ++t) {
        if (idx[t] < 0) continue;
        const int64_t i = idx[t];
        const float pi = __builtin_fmaf(bt, pv[t], uv[t]);
        const float si = __builtin_fmaf(bt, sv[t], wv[t]);
        p[i] = pi;
        s[i] = si;
        d[i] = __builtin_fmaf(a, pi, dv[t]);
        float ui, ri;
        if (invd) {
            ui = __builtin_fmaf(-a, mv[t] * si, uv[t]);
            ri = ui / mv[t];
        } else {
            ui = __builtin_fmaf(-a, si, uv[t]);
            ri = ui;
        }
        u[i] = ui;
        v[0] += (double)ri * (double)ui;
        v[1] += (double)ri * (double)ri;
    }
    const int64_t nb = B.first_block[pid + 1] - B.first_block[pid];
    block_partials<2>(v, lb, nbu, part + M.pbase[pid] + nb);
}

// x += res d for every problem whose inner solve ran this cycle
__global__ void k_bm_update_x(BatchDev B, BMixDev M, int64_t total_blocks, double* __restrict__ x,
                              const float* __restrict__ d, const BMState* st) {
    int pid; int64_t lb, nb;
    locate(B, total_blocks, pid, lb, nb);
    if (st[pid].done) return;
    const int64_t rl = lb * kBlock + threadIdx.x;
    if (rl < B.nrows[pid]) x[B.voff[pid] + rl] += st[pid].m.res * (double)d[M.off32[pid] + rl];
}

}  // namespace rsd

using namespace rsd;

namespace rsd {
// a batch's schedule: none of its fields changes an operation or its order
struct BSched {
    int32_t chunk = 8;   // iterations per launch of the chunk graph
    int32_t unroll = 9;  // row product load batch (4 or 9)
    int32_t order = 0;   // 0: blocks problem by problem, 1: interleaved across problems
    int32_t xcd = 1;     // remap launched blocks XCD by XCD
};
}  // namespace rsd

struct rsd_batch {
    rsd_ctx* ctx = nullptr;
    int32_t np = 0;
    int32_t mode = COL_ABS;  // COL_PAT when every problem has shared offset patterns
    int64_t total_rows = 0, total_blocks = 0, total_nnz = 0;
    rsd::BSched sched;
    std::vector<int64_t> h_voff, h_nrows, h_first;
    DBuf<const int64_t*> d_slice;
    DBuf<const int32_t*> d_rp, d_col, d_diag, d_poff, d_plen;
    DBuf<const uint8_t*> d_pid;
    DBuf<int32_t> d_stride;
    DBuf<const double*> d_val;
    // the problems' FP64 stencil copies, when every problem has one of the
    // same size (st_nr runs of three: 9 hex, 3 quad; 0: the sliced products)
    int32_t st_nr = 0;
    std::vector<const double*> h_val_st;
    DBuf<const double*> d_val_st;
    DBuf<const uint32_t*> d_st_mask;
    DBuf<int32_t> d_st_coff;
    DBuf<int32_t> blk2prob[2], blk2local[2];  // per block order
    DBuf<int64_t> first_block, nrows, voff;
    DBuf<double> r, p, Ap, invd;
    DBuf<BState> st;
    DBuf<double> partials;
    // batched mixed precision, set up on first use
    std::vector<rsd_matrix*> mats;
    int32_t mode32 = -1;  // FP32 column mode of every problem
    int64_t total_t = 0;  // 1024-row tile blocks
    DBuf<Cols> d_c32;
    DBuf<const float*> d_val32;
    DBuf<int32_t> t2prob[2], t2local[2];
    DBuf<int64_t> first_t, pbase, off32;
    int64_t total32 = 0;  // FP32 vector length: the problems' segments and the zeros around them
    DBuf<float> mu, mw, mp, ms, md, minvd;
    DBuf<double> r64, mpart;
    DBuf<BMState> mst;
    BMixDev mdev() const {
        BMixDev M;
        M.c32 = d_c32.p;
        M.val32 = d_val32.p;
        M.t2prob = t2prob[sched.order].p;
        M.t2local = t2local[sched.order].p;
        M.first_t = first_t.p;
        M.pbase = pbase.p;
        M.off32 = off32.p;
        return M;
    }
    BatchDev dev() const {
        BatchDev B;
        B.slice_ptr = d_slice.p;
        B.rp = d_rp.p;
        B.col = d_col.p;
        B.pid = d_pid.p;
        B.poff = d_poff.p;
        B.plen = d_plen.p;
        B.stride = d_stride.p;
        B.val = d_val.p;
        B.diag = d_diag.p;
        B.blk2prob = blk2prob[sched.order].p;
        B.blk2local = blk2local[sched.order].p;
        B.first_block = first_block.p;
        B.nrows = nrows.p;
        B.voff = voff.p;
        B.val_st = d_val_st.p;
        B.st_mask = d_st_mask.p;
        B.st_coff = d_st_coff.p;
        B.xcd = sched.xcd;
        return B;
    }
};

namespace rsd {

// the SpMV kernels for the batch's column mode and load batch
#define RSD_B_SPMV(kern, bt, ...)                                                                                  \
    do {                                                                                                           \
        const dim3 g_((unsigned)(bt)->total_blocks), b_(kBlock);                                                   \
        hipStream_t s_ = (bt)->ctx->stream;                                                                        \
        if ((bt)->mode == COL_PAT) {                                                                               \
            if ((bt)->sched.unroll == 4) hipLaunchKernelGGL((kern<COL_PAT, 4>), g_, b_, 0, s_, __VA_ARGS__);       \
            else hipLaunchKernelGGL((kern<COL_PAT, 9>), g_, b_, 0, s_, __VA_ARGS__);                               \
        } else {                                                                                                   \
            if ((bt)->sched.unroll == 4) hipLaunchKernelGGL((kern<COL_ABS, 4>), g_, b_, 0, s_, __VA_ARGS__);       \
            else hipLaunchKernelGGL((kern<COL_ABS, 9>), g_, b_, 0, s_, __VA_ARGS__);                               \
        }                                                                                                          \
    } while (0)

static void b_launch_iter(rsd_batch* bt, double* x, double tol, int64_t max_iters, int32_t precond) {
    hipStream_t st = bt->ctx->stream;
    const BatchDev B = bt->dev();
    const int64_t nb = bt->total_blocks;
    const double* invd = precond == RSD_PC_JACOBI ? bt->invd.p : nullptr;
    const unsigned np = (unsigned)bt->np;
    if (bt->st_nr == 9)
        hipLaunchKernelGGL(k_b_spmv_pap_st<9>, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, bt->p.p, bt->Ap.p,
                           bt->st.p, bt->partials.p);
    else if (bt->st_nr == 3)
        hipLaunchKernelGGL(k_b_spmv_pap_st<3>, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, bt->p.p, bt->Ap.p,
                           bt->st.p, bt->partials.p);
    else
        RSD_B_SPMV(k_b_spmv_pap, bt, B, nb, bt->p.p, bt->Ap.p, bt->st.p, bt->partials.p);
    hipLaunchKernelGGL(k_b_alpha, dim3(np), dim3(kFold), 0, st, B, bt->partials.p, bt->st.p);
    hipLaunchKernelGGL(k_b_update_r, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, x, bt->p.p, bt->Ap.p, bt->r.p,
                       invd, bt->st.p, bt->partials.p);
    hipLaunchKernelGGL(k_b_beta, dim3(np), dim3(kFold), 0, st, B, bt->partials.p, bt->st.p, tol, max_iters);
    hipLaunchKernelGGL(k_b_update_p, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, bt->p.p, bt->r.p, invd, bt->st.p);
}

// the solve's start: Jacobi, r = b - A x, p, the initial scalars
static void b_launch_init(rsd_batch* bt, const double* b, const double* x, double tol, int64_t max_iters,
                          int32_t precond) {
    hipStream_t st = bt->ctx->stream;
    if (bt->st_nr) {
        // refresh every problem's stencil copy from its values; the copies
        // are allocated once, so their addresses change only the first time
        bool moved = false;
        for (int i = 0; i < bt->np; ++i) {
            reserve_stencil(bt->ctx, bt->mats[i]);
            launch_to_stencil(bt->ctx, bt->mats[i]);
            moved = moved || bt->h_val_st[i] != bt->mats[i]->val_st.p;
            bt->h_val_st[i] = bt->mats[i]->val_st.p;
        }
        if (moved) bt->d_val_st.upload(bt->h_val_st.data(), bt->np, st);
    }
    const BatchDev B = bt->dev();
    const int64_t nb = bt->total_blocks;
    if (precond == RSD_PC_JACOBI)
        hipLaunchKernelGGL(k_b_inv_diag, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, bt->invd.p);
    RSD_B_SPMV(k_b_spmv, bt, B, nb, x, bt->Ap.p);
    hipLaunchKernelGGL(k_b_init, dim3((unsigned)nb), dim3(kBlock), 0, st, B, nb, b, bt->Ap.p, bt->r.p, bt->p.p,
                       precond == RSD_PC_JACOBI ? bt->invd.p : nullptr, bt->partials.p);
    hipLaunchKernelGGL(k_b_init_fold, dim3((unsigned)bt->np), dim3(kFold), 0, st, B, bt->partials.p, bt->st.p, tol,
                       max_iters);
    RSD_HIP(hipGetLastError());
}

struct BGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~BGraph() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
};

// one chunk of iterations captured as a graph
static void b_capture(rsd_batch* bt, double* x, double tol, int64_t max_iters, int32_t precond, BGraph& G) {
    hipStream_t st = bt->ctx->stream;
    RSD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int64_t k = 0; k < bt->sched.chunk; ++k) b_launch_iter(bt, x, tol, max_iters, precond);
    RSD_HIP(hipStreamEndCapture(st, &G.g));
    RSD_HIP(hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0));
}

void batched_cg(rsd_batch* bt, const double* b, double* x, double tol, int64_t max_iters, int32_t precond,
                bool use_graph, std::vector<BState>& out) {
    hipStream_t st = bt->ctx->stream;
    const int64_t chunk = bt->sched.chunk;
    b_launch_init(bt, b, x, tol, max_iters, precond);
    BGraph G;
    if (use_graph) b_capture(bt, x, tol, max_iters, precond, G);
    hipGraphExec_t ge = G.e;
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
        throw;
    }
    (void)hipHostFree(h);
}

// the batched mixed solve's tables and vectors; the FP32 copies are
// refreshed by every solve
static void bm_setup(rsd_batch* bt) {
    rsd_ctx* ctx = bt->ctx;
    hipStream_t st = ctx->stream;
    const int np = bt->np;
    if (bt->mode32 < 0) {
        std::vector<Cols> c32(np);
        std::vector<int64_t> first{0}, pbase(np), off32(np);
        int64_t pb_acc = 0;
        // every problem on the uniform stencil: its FP32 product (bitwise the
        // per-pattern walk of the same storage); else symmetric half storage;
        // else the FP32 copy in the sliced ELL layout walked by absolute
        // columns (bitwise any column mode)
        bool all_uni = true, all_half = true, any_half = false;
        for (int i = 0; i < np; ++i) {
            alloc_f32(ctx, bt->mats[i]);  // builds the half layout when the pattern allows it
            const Sell& S = bt->mats[i]->pb->sell;
            all_uni = all_uni && S.has_uni;
            all_half = all_half && S.has_half;
            any_half = any_half || S.has_half;
        }
        RSD_CHECK(all_half || !any_half, "batched mixed CG: the problems' FP32 layouts differ");
        bt->mode32 = all_uni ? COL_UNI : all_half ? COL_HALF : COL_ABS;
        // FP32 vectors: problem i's rows from off32[i], its 256-row blocks,
        // and the zeros its uniform-stencil product reads before and after
        // them (none for the other layouts), segment starts on 64 floats
        int64_t acc = 0;
        auto up64 = [](int64_t v) { return (v + 63) / 64 * 64; };
        for (int i = 0; i < np; ++i) {
            const Sell& S = bt->mats[i]->pb->sell;
            const bool uni = bt->mode32 == COL_UNI;
            acc += up64(uni ? S.uni_lo : 0);
            off32[i] = acc;
            acc += std::max<int64_t>(1, nblocks(S.n_rows)) * kBlock + up64(uni ? S.uni_hi : 0);
        }
        bt->total32 = acc;
        for (int i = 0; i < np; ++i) {
            rsd_matrix* A = bt->mats[i];
            const Sell& S = A->pb->sell;
            c32[i] = S.cols32();
            const int64_t nb = std::max<int64_t>(1, nblocks(S.n_rows)), nbu = mixed_blocks(S.n_rows);
            first.push_back(first.back() + nbu);
            pbase[i] = pb_acc;
            pb_acc += nb + 2 * nbu;
        }
        bt->total_t = first.back();
        std::vector<int32_t> t2p[2], t2l[2];
        for (int i = 0; i < np; ++i)
            for (int64_t b = first[i]; b < first[i + 1]; ++b) {
                t2p[0].push_back(i);
                t2l[0].push_back((int32_t)(b - first[i]));
            }
        for (int64_t k = 0; (int64_t)t2p[1].size() < bt->total_t; ++k)
            for (int i = 0; i < np; ++i)
                if (k < first[i + 1] - first[i]) {
                    t2p[1].push_back(i);
                    t2l[1].push_back((int32_t)k);
                }
        bt->d_c32.upload(c32.data(), np, st);
        for (int o = 0; o < 2; ++o) {
            bt->t2prob[o].upload(t2p[o].data(), bt->total_t, st);
            bt->t2local[o].upload(t2l[o].data(), bt->total_t, st);
        }
        bt->first_t.upload(first.data(), np + 1, st);
        bt->pbase.upload(pbase.data(), np, st);
        bt->off32.upload(off32.data(), np, st);
        RSD_HIP(hipStreamSynchronize(st));  // the tables are locals
        const int64_t n = bt->total32;
        for (DBuf<float>* v : {&bt->mu, &bt->mw, &bt->mp, &bt->ms, &bt->md, &bt->minvd}) {
            v->alloc(n);
            launch_fill_f32(ctx, n, v->p, 0.0f);
        }
        bt->r64.alloc(bt->total_rows);
        bt->mpart.alloc(pb_acc);
        bt->mst.alloc(np);
    }
    std::vector<const float*> v32(np);
    for (int i = 0; i < np; ++i) {
        matrix_to_f32(ctx, bt->mats[i], true);
        v32[i] = bt->mats[i]->val32.p;
    }
    bt->d_val32.upload(v32.data(), np, st);
    // the FP32 conversion's symmetry check, as the standalone solve reports it
    for (int i = 0; i < np; ++i) {
        const rsd_matrix* A = bt->mats[i];
        if (!A->pb->sell.has_half || !A->asym.n) continue;
        int32_t flag = 0;
        RSD_HIP(hipMemcpyAsync(&flag, A->asym.p, sizeof flag, hipMemcpyDeviceToHost, st));
        RSD_HIP(hipStreamSynchronize(st));
        if (flag)
            throw Error(RSD_ERR_ARG, "batched mixed CG: problem " + std::to_string(i) + "'s matrix is not symmetric "
                                     "(a value differs from its mirror by more than 1e-10 of its row's largest entry)");
    }
}

#define RSD_BM_SPMV32(bt, grid, ...)                                                                               \
    do {                                                                                                           \
        hipStream_t s_ = (bt)->ctx->stream;                                                                        \
        if ((bt)->mode32 == COL_UNI) hipLaunchKernelGGL((k_bm_spmv_wu<COL_UNI>), grid, dim3(kBlock), 0, s_, __VA_ARGS__); \
        else if ((bt)->mode32 == COL_HALF) hipLaunchKernelGGL((k_bm_spmv_wu<COL_HALF>), grid, dim3(kBlock), 0, s_, __VA_ARGS__); \
        else hipLaunchKernelGGL((k_bm_spmv_wu<COL_ABS>), grid, dim3(kBlock), 0, s_, __VA_ARGS__);                \
    } while (0)

static void bm_launch_inner_iter(rsd_batch* bt, int32_t precond) {
    hipStream_t st = bt->ctx->stream;
    const BatchDev B = bt->dev();
    const BMixDev M = bt->mdev();
    const float* invd = precond == RSD_PC_JACOBI ? bt->minvd.p : nullptr;
    RSD_BM_SPMV32(bt, dim3((unsigned)bt->total_blocks), B, M, bt->total_blocks, bt->mu.p, bt->mw.p, bt->mst.p,
                  bt->mpart.p);
    hipLaunchKernelGGL(k_bm_scalars, dim3((unsigned)bt->np), dim3(kFold), 0, st, B, M, bt->mpart.p, bt->mst.p);
    hipLaunchKernelGGL(k_bm_update, dim3((unsigned)bt->total_t), dim3(kBlock), 0, st, B, M, bt->total_t, bt->mp.p,
                       bt->ms.p, bt->md.p, bt->mu.p, bt->mw.p, invd, bt->mst.p, bt->mpart.p);
}

void batched_mixed(rsd_batch* bt, const double* b, double* x, double tol, int64_t max_iters, int32_t precond,
                   double inner_floor, std::vector<BMState>& out) {
    hipStream_t st = bt->ctx->stream;
    bm_setup(bt);
    const BatchDev B = bt->dev();
    const BMixDev M = bt->mdev();
    const int np = bt->np;
    const int64_t nb = bt->total_blocks;
    if (precond == RSD_PC_JACOBI)
        hipLaunchKernelGGL(k_bm_inv_diag, dim3((unsigned)nb), dim3(kBlock), 0, st, B, M, nb, bt->minvd.p);
    const float* invd = precond == RSD_PC_JACOBI ? bt->minvd.p : nullptr;
    BMState* h = nullptr;
    RSD_HIP(hipHostMalloc((void**)&h, sizeof(BMState) * (size_t)np, hipHostMallocDefault));
    std::memset(h, 0, sizeof(BMState) * (size_t)np);
    for (int i = 0; i < np; ++i) h[i].m.in_max = max_iters;
    BGraph G;
    try {
        RSD_HIP(hipMemcpyAsync(bt->mst.p, h, sizeof(BMState) * (size_t)np, hipMemcpyHostToDevice, st));
        // one chunk of inner iterations for the whole batch, as a graph
        RSD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int64_t k = 0; k < bt->sched.chunk; ++k) bm_launch_inner_iter(bt, precond);
        RSD_HIP(hipStreamEndCapture(st, &G.g));
        RSD_HIP(hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0));
        auto read_states = [&] {
            RSD_HIP(hipMemcpyAsync(h, bt->mst.p, sizeof(BMState) * (size_t)np, hipMemcpyDeviceToHost, st));
            RSD_HIP(hipStreamSynchronize(st));
        };
        for (int cycle = 0;; ++cycle) {
            RSD_CHECK(cycle <= kMixedMaxCycles + 1, "batched mixed CG: outer loop did not end");
            if (bt->mode == COL_PAT) {
                if (bt->sched.unroll == 4)
                    hipLaunchKernelGGL((k_bm_residual<COL_PAT, 4>), dim3((unsigned)nb), dim3(kBlock), 0, st, B, M, nb,
                                       b, x, bt->r64.p, bt->mst.p, bt->mpart.p);
                else
                    hipLaunchKernelGGL((k_bm_residual<COL_PAT, 9>), dim3((unsigned)nb), dim3(kBlock), 0, st, B, M, nb,
                                       b, x, bt->r64.p, bt->mst.p, bt->mpart.p);
            } else {
                hipLaunchKernelGGL((k_bm_residual<COL_ABS, 9>), dim3((unsigned)nb), dim3(kBlock), 0, st, B, M, nb, b,
                                   x, bt->r64.p, bt->mst.p, bt->mpart.p);
            }
            hipLaunchKernelGGL(k_bm_outer, dim3((unsigned)np), dim3(kFold), 0, st, B, M, bt->mpart.p, bt->mst.p, tol,
                               max_iters, inner_floor);
            read_states();
            bool all_done = true;
            for (int i = 0; i < np; ++i) all_done = all_done && h[i].done;
            if (all_done) break;
            hipLaunchKernelGGL(k_bm_init, dim3((unsigned)bt->total_t), dim3(kBlock), 0, st, B, M, bt->total_t,
                               bt->r64.p, bt->mu.p, bt->mp.p, bt->ms.p, bt->md.p, invd, bt->mst.p, bt->mpart.p);
            int64_t launched = 0;
            for (;;) {
                RSD_HIP(hipGraphLaunch(G.e, st));
                launched += bt->sched.chunk;
                read_states();
                bool stopped = true;
                for (int i = 0; i < np; ++i) stopped = stopped && h[i].m.stop;
                if (stopped) break;
                RSD_CHECK(launched <= max_iters + 2 * bt->sched.chunk, "batched mixed CG: inner solve did not stop");
            }
            hipLaunchKernelGGL(k_bm_update_x, dim3((unsigned)nb), dim3(kBlock), 0, st, B, M, nb, x, bt->md.p, bt->mst.p);
            RSD_HIP(hipGetLastError());
        }
        out.assign(h, h + np);
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
}

// ------------------------------------------------------------------------
// autotuner
// ------------------------------------------------------------------------
// the cache key: the device architecture and what sets the batch's cost
static std::string batch_key(const rsd_batch* bt, int32_t precond) {
    hipDeviceProp_t prop;
    int dev = 0;
    RSD_HIP(hipGetDevice(&dev));
    RSD_HIP(hipGetDeviceProperties(&prop, dev));
    std::ostringstream k;
    k << prop.gcnArchName << "/np=" << bt->np << "/rows=" << bt->total_rows << "/nnz=" << bt->total_nnz
      << "/cols=" << (bt->mode == COL_PAT ? "pat" : "abs") << "/pc=" << (precond == RSD_PC_JACOBI ? "jacobi" : "identity");
    return k.str();
}

struct TuneEntry {
    BSched s;
    double us = 0.0;
};

// The cache file holds one entry per line:
//   "key": {"chunk": C, "unroll": U, "order": O, "xcd": X, "us_per_iter": T},
// between a line "{" and a line "}" (a JSON object).
// the ranges rsd_batch_set_schedule accepts
static bool valid_schedule(const BSched& s) {
    return s.chunk >= 1 && s.chunk <= 256 && (s.unroll == 4 || s.unroll == 9) && (s.order == 0 || s.order == 1) &&
           (s.xcd == 0 || s.xcd == 1);
}

// entries outside the accepted ranges are dropped, so their keys are tuned again
static std::map<std::string, TuneEntry> read_cache(const std::string& path) {
    std::map<std::string, TuneEntry> m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        const size_t q0 = line.find('"'), q1 = q0 == std::string::npos ? q0 : line.find('"', q0 + 1);
        const size_t br = q1 == std::string::npos ? q1 : line.find('{', q1);
        if (br == std::string::npos) continue;
        TuneEntry e;
        if (std::sscanf(line.c_str() + br, "{\"chunk\": %d, \"unroll\": %d, \"order\": %d, \"xcd\": %d, \"us_per_iter\": %lf",
                        &e.s.chunk, &e.s.unroll, &e.s.order, &e.s.xcd, &e.us) == 5 &&
            valid_schedule(e.s))
            m[line.substr(q0 + 1, q1 - q0 - 1)] = e;
    }
    return m;
}

static void write_cache(const std::string& path, const std::map<std::string, TuneEntry>& m) {
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp);
        RSD_CHECK(f.good(), "rsd_batch_autotune: cannot write the cache file");
        f << "{\n";
        size_t i = 0;
        for (const auto& [k, e] : m) {
            char buf[256];
            std::snprintf(buf, sizeof buf, "{\"chunk\": %d, \"unroll\": %d, \"order\": %d, \"xcd\": %d, \"us_per_iter\": %.3f}",
                          e.s.chunk, e.s.unroll, e.s.order, e.s.xcd, e.us);
            f << "\"" << k << "\": " << buf << (++i < m.size() ? ",\n" : "\n");
        }
        f << "}\n";
    }
    RSD_CHECK(std::rename(tmp.c_str(), path.c_str()) == 0, "rsd_batch_autotune: cannot replace the cache file");
}

// time one schedule: the real loop (graph launch of a chunk, state read,
// synchronisation) over about `iters` iterations of a problem that never
// converges (tol 0), from b = 1, x = 0 in scratch vectors
static double time_schedule(rsd_batch* bt, int32_t precond, double* b, double* x, int64_t iters) {
    hipStream_t st = bt->ctx->stream;
    const int64_t big = int64_t(1) << 40;
    launch_fill(bt->ctx, bt->total_rows, x, 0.0);
    b_launch_init(bt, b, x, 0.0, big, precond);
    BGraph G;
    b_capture(bt, x, 0.0, big, precond, G);
    BState* h = nullptr;
    RSD_HIP(hipHostMalloc((void**)&h, sizeof(BState) * (size_t)bt->np, hipHostMallocDefault));
    double us = 0.0;
    try {
        auto round = [&] {
            RSD_HIP(hipGraphLaunch(G.e, st));
            RSD_HIP(hipMemcpyAsync(h, bt->st.p, sizeof(BState) * (size_t)bt->np, hipMemcpyDeviceToHost, st));
            RSD_HIP(hipStreamSynchronize(st));
        };
        round();  // warm-up
        const int64_t rounds = std::max<int64_t>(1, (iters + bt->sched.chunk - 1) / bt->sched.chunk);
        const auto t0 = std::chrono::steady_clock::now();
        for (int64_t k = 0; k < rounds; ++k) round();
        const auto t1 = std::chrono::steady_clock::now();
        us = std::chrono::duration<double, std::micro>(t1 - t0).count() / (double)(rounds * bt->sched.chunk);
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
    return us;
}

}  // namespace rsd

extern "C" {

int rsd_batch_create(rsd_ctx* ctx, int32_t n_problems, rsd_matrix* const* mats, rsd_batch** out) {
    return guard([&] {
        RSD_CHECK(ctx && out && n_problems > 0 && mats, "rsd_batch_create: bad argument");
        auto bt = std::make_unique<rsd_batch>();
        bt->ctx = ctx;
        bt->np = n_problems;
        bt->mats.assign(mats, mats + n_problems);
        std::vector<const int64_t*> sl(n_problems);
        std::vector<const int32_t*> rp(n_problems), col(n_problems), dg(n_problems), poff(n_problems), plen(n_problems);
        std::vector<const uint8_t*> pid(n_problems);
        std::vector<int32_t> stride(n_problems, 0);
        std::vector<const double*> val(n_problems);
        bool all_pat = true;
        bt->h_voff.assign(n_problems + 1, 0);
        bt->h_first.assign(n_problems + 1, 0);
        bt->h_nrows.assign(n_problems, 0);
        for (int i = 0; i < n_problems; ++i) {
            RSD_CHECK(mats[i], "rsd_batch_create: null matrix");
            RSD_LIVE_MATRIX(mats[i]);
            RSD_CHECK(mats[i]->pb->ctx == ctx, "rsd_batch_create: matrix on another context");
            const Sell& S = mats[i]->pb->sell;
            sl[i] = S.slice_ptr.p;
            rp[i] = S.row_ptr.p;
            col[i] = S.col.p;
            dg[i] = S.diag.p;
            pid[i] = S.pat_id.p;
            poff[i] = S.pat_off.p;
            plen[i] = S.pat_len.p;
            stride[i] = S.pat_stride;
            all_pat = all_pat && S.has_pat;
            bt->total_nnz += mats[i]->pb->nnz;
            val[i] = mats[i]->val.p;
            const int64_t nbk = std::max<int64_t>(1, nblocks(S.n_rows));
            bt->h_nrows[i] = S.n_rows;
            bt->h_first[i + 1] = bt->h_first[i] + nbk;
            bt->h_voff[i + 1] = bt->h_voff[i] + nbk * kBlock;
        }
        bt->total_blocks = bt->h_first[n_problems];
        bt->total_rows = bt->h_voff[n_problems];
        bt->mode = all_pat ? COL_PAT : COL_ABS;
        // every problem with a stencil copy of the same size: the CG products
        // read those (bitwise the sliced products), refreshed by each solve
        bool all_st = true;
        for (int i = 0; i < n_problems; ++i) {
            const Sell& S = mats[i]->pb->sell;
            all_st = all_st && S.has_st && S.n_rows > 0 && S.st_np == mats[0]->pb->sell.st_np;
        }
        if (all_st) {
            hipStream_t st = ctx->stream;
            bt->st_nr = mats[0]->pb->sell.st_np / 3;
            std::vector<const uint32_t*> smask(n_problems);
            std::vector<int32_t> scoff((size_t)n_problems * kStRuns, 0);
            for (int i = 0; i < n_problems; ++i) {
                const StCols c = mats[i]->pb->sell.st_cols();
                smask[i] = c.mask;
                for (int k = 0; k < bt->st_nr; ++k) scoff[(size_t)i * kStRuns + k] = c.coff[k];
            }
            bt->d_st_mask.upload(smask.data(), n_problems, st);
            bt->d_st_coff.upload(scoff.data(), (int64_t)scoff.size(), st);
            bt->h_val_st.assign(n_problems, nullptr);
            bt->d_val_st.upload(bt->h_val_st.data(), n_problems, st);
            RSD_HIP(hipStreamSynchronize(st));  // the tables are locals
        }
        // block orders: problem by problem, and interleaved (block k of every
        // problem that has one, then block k + 1, ...)
        std::vector<int32_t> b2p[2], b2l[2];
        for (int i = 0; i < n_problems; ++i)
            for (int64_t b = bt->h_first[i]; b < bt->h_first[i + 1]; ++b) {
                b2p[0].push_back(i);
                b2l[0].push_back((int32_t)(b - bt->h_first[i]));
            }
        for (int64_t k = 0; (int64_t)b2p[1].size() < bt->total_blocks; ++k)
            for (int i = 0; i < n_problems; ++i)
                if (k < bt->h_first[i + 1] - bt->h_first[i]) {
                    b2p[1].push_back(i);
                    b2l[1].push_back((int32_t)k);
                }
        hipStream_t st = ctx->stream;
        bt->d_slice.upload(sl.data(), n_problems, st);
        bt->d_rp.upload(rp.data(), n_problems, st);
        bt->d_col.upload(col.data(), n_problems, st);
        bt->d_diag.upload(dg.data(), n_problems, st);
        bt->d_val.upload(val.data(), n_problems, st);
        for (int o = 0; o < 2; ++o) {
            bt->blk2prob[o].upload(b2p[o].data(), bt->total_blocks, st);
            bt->blk2local[o].upload(b2l[o].data(), bt->total_blocks, st);
        }
        bt->d_pid.upload(pid.data(), n_problems, st);
        bt->d_poff.upload(poff.data(), n_problems, st);
        bt->d_plen.upload(plen.data(), n_problems, st);
        bt->d_stride.upload(stride.data(), n_problems, st);
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
        batched_cg(bt, b, x, tol, max_iters, precond, use_graph != 0, s);
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

int rsd_cg_solve_batched_mixed(rsd_ctx* ctx, rsd_batch* bt, const double* b, double* x, double tol,
                               int64_t max_iters, int32_t precond, double inner_floor, rsd_report* reps,
                               int64_t* outer_cycles) {
    return guard([&] {
        RSD_CHECK(ctx && bt && b && x && ctx == bt->ctx, "rsd_cg_solve_batched_mixed: bad argument");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        std::vector<BMState> s;
        batched_mixed(bt, b, x, tol, max_iters, precond, inner_floor > 0 ? inner_floor : kMixedDefaultInnerFloor, s);
        for (int i = 0; i < bt->np; ++i) {
            const MixedState& m = s[i].m;
            if (reps) {
                reps[i].converged = m.converged;
                reps[i].iterations = m.total_inner;
                reps[i].initial_residual = m.res0;
                reps[i].final_residual = m.res;
                reps[i].final_residual_rel = m.res / (m.res0 + 1e-50);
            }
            if (outer_cycles) outer_cycles[i] = m.cycles;
        }
    });
}

int rsd_batch_set_schedule(rsd_batch* bt, const int32_t* sched) {
    return guard([&] {
        RSD_CHECK(bt && sched, "rsd_batch_set_schedule: null argument");
        RSD_CHECK(sched[0] >= 1 && sched[0] <= 256, "chunk must be 1..256");
        RSD_CHECK(sched[1] == 4 || sched[1] == 9, "unroll must be 4 or 9");
        RSD_CHECK(sched[2] == 0 || sched[2] == 1, "order must be 0 or 1");
        RSD_CHECK(sched[3] == 0 || sched[3] == 1, "xcd must be 0 or 1");
        bt->sched = BSched{sched[0], sched[1], sched[2], sched[3]};
    });
}

int rsd_batch_get_schedule(rsd_batch* bt, int32_t* sched) {
    return guard([&] {
        RSD_CHECK(bt && sched, "rsd_batch_get_schedule: null argument");
        sched[0] = bt->sched.chunk;
        sched[1] = bt->sched.unroll;
        sched[2] = bt->sched.order;
        sched[3] = bt->sched.xcd;
    });
}

int rsd_batch_autotune(rsd_ctx* ctx, rsd_batch* bt, int32_t precond, const char* cache_path, int64_t iters,
                       int32_t* sched, double* us_per_iter, int32_t* from_cache) {
    return guard([&] {
        RSD_CHECK(ctx && bt && ctx == bt->ctx, "rsd_batch_autotune: bad argument");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        const std::string key = batch_key(bt, precond);
        std::map<std::string, TuneEntry> cache;
        if (cache_path && *cache_path) cache = read_cache(cache_path);
        TuneEntry best;
        bool cached = false;
        if (auto it = cache.find(key); it != cache.end()) {
            best = it->second;
            cached = true;
        } else {
            DBuf<double> b, x;
            b.alloc(bt->total_rows);
            x.alloc(bt->total_rows);
            launch_fill(ctx, bt->total_rows, b.p, 1.0);
            best.us = 1e300;
            for (int32_t chunk : {4, 8, 16, 32})
                for (int32_t unroll : {4, 9})
                    for (int32_t order : {0, 1})
                        for (int32_t xcd : {0, 1}) {
                            bt->sched = BSched{chunk, unroll, order, xcd};
                            const double us = time_schedule(bt, precond, b.p, x.p, std::max<int64_t>(iters, 1));
                            if (us < best.us) best = TuneEntry{bt->sched, us};
                        }
            RSD_HIP(hipStreamSynchronize(ctx->stream));
            if (cache_path && *cache_path) {
                cache[key] = best;
                write_cache(cache_path, cache);
            }
        }
        bt->sched = best.s;
        if (sched) {
            sched[0] = best.s.chunk;
            sched[1] = best.s.unroll;
            sched[2] = best.s.order;
            sched[3] = best.s.xcd;
        }
        if (us_per_iter) *us_per_iter = best.us;
        if (from_cache) *from_cache = cached ? 1 : 0;
    });
}

}  // extern "C"
