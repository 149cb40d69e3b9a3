// rsd_mixed.hip -- mixed-precision CG: FP32 bulk, FP64 refinement.
//
// Iterative refinement around an FP32 preconditioned CG:
//   outer (FP64):  r = b - A x with the FP64 matrix, ||r|| in FP64;
//                  stop when ||r|| / (||r_0|| + 1e-50) < tol, the same
//                  relative criterion as cg/Solver.tpp:45-48;
//   inner (FP32):  solve A d = r / ||r|| with FP32 values, FP32 vectors and
//                  FP32-accumulated SpMV rows, by the Chronopoulos-Gear
//                  form of PCG (one fused reduction of three dot products
//                  per iteration, accumulated in FP64):
//                    u = M r, w = A u, delta = w.u, gamma = r.u,
//                    beta = gamma / gamma', alpha = gamma / (delta - beta gamma / alpha'),
//                    p = u + beta p, s = w + beta s, d += alpha p, r -= alpha s
//                  (r kept implicit as u / invd);
//   update (FP64): x += ||r|| d.
// The FP32 copy of a structurally symmetric matrix whose rows share offset
// patterns keeps each symmetric pair's value once (the upper one; the
// conversion checks that the FP64 matrix is symmetric to 1e-10 of each row's
// largest entry, and the solve refuses one that is not): per inner iteration
// the matrix stream is about 2 bytes per nonzero instead of the FP64
// solver's 8, with no index stream.  An iteration is three launches: the
// SpMV with w.u, the one-block fold of the three dot products, and the
// vector update, which also forms r.u and r.r of the new residual for the
// next iteration.
#include "rsd_linalg.hpp"
#include "rsd_mixed.hpp"
#include "rsd_rowdot.hpp"

#include <algorithm>
#include <cmath>

namespace rsd {

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_set_flag(int32_t* f, int32_t v) {
    if (threadIdx.x == 0) *f = v;
}

// The half layout's FP32 upper values from the FP64 SELL values (one thread
// per row): upper entry j of row r (pattern q) goes to r + dst[q][j] (the
// m-th upper entry to m * pad + r, or in the uniform stencil storage to its
// offset's plane).  With CHECK, the row's values left of the diagonal are also
// compared with the FP64 values stored for the mirrored entries, and *asym
// is set when one differs by more than 1e-10 of the row's largest magnitude
// (an assembled heat matrix is symmetric to rounding).
template<bool CHECK>
__global__ __launch_bounds__(kBlock) void k_to_f32_half(int64_t n, Cols H, const int32_t* __restrict__ mir,
                                                        const int32_t* __restrict__ low,
                                                        const int32_t* __restrict__ dst,
                                                        const double* __restrict__ val, float* __restrict__ vh,
                                                        int32_t* __restrict__ asym) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int q = H.pid[r];
    const int len = H.plen[q], lo = low[q];
    const int64_t base = H.slice_ptr[r >> 6] + (r & 63);
    const int32_t* pd = dst + q * H.stride;
    for (int j = lo; j < len; ++j) vh[(uint32_t)r + (uint32_t)pd[j]] = (float)val[base + (int64_t)j * kSlice];
    if constexpr (CHECK) {
        const int32_t* po = H.poff + q * H.stride;
        const int32_t* pm = mir + q * H.stride;
        double amax = 0.0, dmax = 0.0;
        for (int j = 0; j < len; ++j) amax = fmax(amax, fabs(val[base + (int64_t)j * kSlice]));
        for (int j = 0; j < lo; ++j) {
            const double a = val[base + (int64_t)j * kSlice];
            const int64_t c = r + po[j];
            const int lc = low[H.pid[c]];
            const double b = val[H.slice_ptr[c >> 6] + (int64_t)(lc + pm[j]) * kSlice + (c & 63)];
            dmax = fmax(dmax, fabs(a - b));
        }
        if (dmax > 1e-10 * amax) *asym = 1;
    }
}

template<int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_f32(int64_t n, int64_t nb, Cols C, const float* __restrict__ val,
                                                     const float* __restrict__ x, float* __restrict__ y) {
    const int64_t r = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    if (r < n) y[r] = row_dot_f32<MODE>(r, C, val, x);
}

// outer residual r = b - A x (FP64), block sums of ||r||^2
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_mx_residual(int64_t n, int64_t nb, Cols C, const double* __restrict__ val,
                                                        const double* __restrict__ b, const double* __restrict__ x,
                                                        double* __restrict__ r64, double* partials,
                                                        const MixedState* st) {
    if (st->done) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t i = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const double ri = b[i] + -1.0 * row_dot_f64<MODE>(i, C, val, x);
        r64[i] = ri;
        v[0] = ri * ri;
    }
    block_partials<1>(v, lb, nb, partials);
}

// the first residual from a given A x: b + -1.0 (A x), the row product's
// rounding sequence with the product already formed
__global__ __launch_bounds__(kBlock) void k_mx_residual_ax(int64_t n, int64_t nb, const double* __restrict__ b,
                                                           const double* __restrict__ ax, double* __restrict__ r64,
                                                           double* partials, const MixedState* st) {
    if (st->done) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t i = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const double ri = b[i] + -1.0 * ax[i];
        r64[i] = ri;
        v[0] = ri * ri;
    }
    block_partials<1>(v, lb, nb, partials);
}

// the outer loop's step, on the device: ||r||, then (in the host solver's
// double arithmetic) convergence, the caps, or the next inner solve's
// tolerance max(floor, 0.5 tol ||r_0|| / ||r||) capped at 0.5 and its cap
__global__ __launch_bounds__(kFold) void k_mx_outer(int64_t nb, const double* __restrict__ partials,
                                                    const int32_t* __restrict__ asym, MixedState* st, double tol,
                                                    int64_t max_iters, double inner_floor) {
    if (st->done) return;
    double out[1];
    fold_partials<1>(nb, partials, out);
    if (threadIdx.x != 0) return;
    const double res = sqrt(out[0]);
    if (st->cycles == 0) st->res0 = res;
    st->res = res;
    st->asym = asym ? *asym : 0;
    const double rel = res / (st->res0 + 1e-50);
    if (st->asym || rel < tol || st->cycles >= kMixedMaxCycles || st->total_inner >= max_iters || !(rel == rel)) {
        st->converged = !st->asym && rel < tol;
        st->done = 1;
        st->stop = 1;
        return;
    }
    double it = 0.5 * tol * (st->res0 + 1e-50) / res;
    it = (inner_floor < it) ? it : inner_floor;  // max(inner_floor, it)
    it = (it < 0.5) ? it : 0.5;                  // min(0.5, it)
    st->in_tol = it;
    st->in_max = max_iters - st->total_inner > 1 ? max_iters - st->total_inner : 1;
    st->cycles = st->cycles + 1;
    st->it = 0;
    st->stop = 0;
}

// inner set-up: r = r64 / ||r||, u = M r, p = s = d = 0; block sums of r.u,
// r.r.  Only u is stored: the residual of row i is u_i / invd_i.
__global__ __launch_bounds__(kBlock) void k_mx_init(int64_t n, int64_t nb, const double* __restrict__ r64,
                                                    float* __restrict__ u, float* __restrict__ p,
                                                    float* __restrict__ s, float* __restrict__ d,
                                                    const float* __restrict__ invd, const MixedState* st,
                                                    double* P12) {
    if (st->done) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const double sc = st->res > 0.0 ? 1.0 / st->res : 0.0;
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int t = 0; t < kMxTiles; ++t) {
        const int64_t i = (lb * kMxTiles + t) * kBlock + threadIdx.x;
        if (i < n) {
            const float ri = (float)(r64[i] * sc);
            const float ui = invd ? ri * invd[i] : ri;
            u[i] = ui;
            p[i] = 0.0f;
            s[i] = 0.0f;
            d[i] = 0.0f;
            v[0] += (double)ri * (double)ui;
            v[1] += (double)ri * (double)ri;
        }
    }
    block_partials<2>(v, lb, nb, P12);
}

// w = A u, block sums of w.u (one row per thread)
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_mx_spmv_wu(int64_t n, int64_t nb, Cols C, const float* __restrict__ val,
                                                       const float* __restrict__ u, float* __restrict__ w,
                                                       const MixedState* st, double* P0) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t i = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const float wi = row_dot_f32<MODE>(i, C, val, u);
        w[i] = wi;
        v[0] = (double)wi * (double)u[i];
    }
    block_partials<1>(v, lb, nb, P0);
}

// fold delta = w.u (nb block sums at P0), gamma = r.u and ||r||^2 (nbu
// block sums each, at P12 and P12 + nbu): convergence, alpha, beta
__global__ __launch_bounds__(kFold) void k_mx_scalars(int64_t nb, const double* __restrict__ P0, int64_t nbu,
                                                      const double* __restrict__ P12, MixedState* st) {
    if (st->stop) return;
    const int64_t k = st->it;
    const double gam_prev = st->gam, alp_prev = st->alp, in_res0 = st->in_res0, in_tol = st->in_tol;
    const int64_t in_max = st->in_max;
    double o[3];
    const double* pk[3] = {P0, P12, P12 + nbu};
    const int64_t nbk[3] = {nb, nbu, nbu};
    fold_partials<3>(pk, nbk, o);
    if (threadIdx.x != 0) return;
    const double del = o[0], gam = o[1], res = sqrt(o[2]);
    const double r0 = k == 0 ? res : in_res0;
    st->total_inner += 1;
    if (k == 0) st->in_res0 = res;
    if (res / (r0 + 1e-50) < in_tol || !(res == res) || k >= in_max) {
        st->stop = 1;
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
    st->alpha = alpha;
    st->beta = beta;
    st->gam = gam;
    st->alp = alpha;
    st->it = k + 1;
}

// p = u + beta p, s = w + beta s, d += alpha p, and the preconditioned
// residual u = M r of r -= alpha s as u -= alpha (M s), the residual itself
// being u / invd (Jacobi) or u; block sums of the new r.u, r.r.  Keeping r
// implicit saves one read and one write of a vector per iteration.  A
// thread's loads for all its tiles are issued before the arithmetic.
__global__ __launch_bounds__(kBlock) void k_mx_update(int64_t n, int64_t nb, float* __restrict__ p,
                                                      float* __restrict__ s, float* __restrict__ d,
                                                      float* __restrict__ u, const float* __restrict__ w,
                                                      const float* __restrict__ invd, const MixedState* st,
                                                      double* P12) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const float a = (float)st->alpha, bt = (float)st->beta;
    constexpr int T = kMxTiles;
    int64_t idx[T];
    float pv[T], sv[T], dv[T], uv[T], wv[T], mv[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        idx[t] = (lb * T + t) * kBlock + threadIdx.x;
        const bool in = idx[t] < n;
        pv[t] = in ? p[idx[t]] : 0.0f;
        sv[t] = in ? s[idx[t]] : 0.0f;
        dv[t] = in ? d[idx[t]] : 0.0f;
        uv[t] = in ? u[idx[t]] : 0.0f;
        wv[t] = in ? w[idx[t]] : 0.0f;
        mv[t] = (in && invd) ? invd[idx[t]] : 1.0f;
    }
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int t = 0; t < T; ++t) {
        if (idx[t] >= n) continue;
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
    block_partials<2>(v, lb, nb, P12);
}

__global__ void k_mx_update_x(int64_t n, double* __restrict__ x, const float* __restrict__ d, const MixedState* st) {
    if (st->done) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] += st->res * (double)d[i];
}

__global__ void k_inv_diag_f32(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ val,
                               float* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t s = diag[i];
    const double dd = s >= 0 ? val[s] : 0.0;
    inv[i] = (float)((fabs(dd) > 1e-300) ? 1.0 / dd : 1.0);
}

// ------------------------------------------------------------------------
void alloc_f32(rsd_ctx* ctx, rsd_matrix* A) {
    build_half(A->pb);
    const Sell& S = A->pb->sell;
    if (A->val32.n == S.slots32()) return;
    A->val32.alloc(S.slots32());
    // the uniform stencil storage holds zeros where a row lacks an entry:
    // the conversion writes only the entries that exist
    if (S.has_uni) launch_fill_f32(ctx, A->val32.n, A->val32.p, 0.0f);
}

void matrix_to_f32(rsd_ctx* ctx, rsd_matrix* A, bool check_symmetry) {
    alloc_f32(ctx, A);
    const Sell& S = A->pb->sell;
    if (S.has_half) {
        if (!A->asym.n) A->asym.alloc(1);
        hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, ctx->stream, A->asym.p, 0);
        if (S.n_rows) {
            const dim3 grid((unsigned)nblocks(S.n_rows));
            if (check_symmetry)
                hipLaunchKernelGGL(k_to_f32_half<true>, grid, dim3(kBlock), 0, ctx->stream, S.n_rows, S.cols32(),
                                   S.half_mir.p, S.half_low.p, S.half_dst.p, A->val.p, A->val32.p, A->asym.p);
            else
                hipLaunchKernelGGL(k_to_f32_half<false>, grid, dim3(kBlock), 0, ctx->stream, S.n_rows, S.cols32(),
                                   S.half_mir.p, S.half_low.p, S.half_dst.p, A->val.p, A->val32.p, A->asym.p);
        }
    } else if (A->val.n) {
        hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nblocks(A->val.n)), dim3(kBlock), 0, ctx->stream, A->val.n,
                           A->val.p, A->val32.p);
    }
}

// padded: x is readable (and finite) from x - uni_lo to x + n + uni_hi
void launch_spmv_f32(rsd_ctx* ctx, const rsd_matrix* A, const float* x, float* y, bool padded) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = nblocks(n);
    if (!n) return;
    RSD_LAUNCH_COLS32(padded ? S.mode_padded() : S.mode32(), k_spmv_f32, dim3((unsigned)nb), dim3(kBlock),
                      ctx->stream, n, nb, S.cols32(), A->val32.p, x, y);
}

MixedWork mixed_workspace(rsd_ctx* ctx, int64_t n, const Sell* S) {
    MixedWork w;
    const size_t f = sizeof(float) * (size_t)std::max<int64_t>(n, 1), dd = sizeof(double) * (size_t)std::max<int64_t>(n, 1);
    w.r64 = (double*)ctx->ws("mx.r64", dd);
    // u, the SpMV operand, carries the uniform stencil's padding
    w.lo = S && S->has_uni ? S->uni_lo : 0;
    w.hi = S && S->has_uni ? S->uni_hi : 0;
    w.ubase = (float*)ctx->ws("mx.u", f + sizeof(float) * (size_t)(w.lo + w.hi));
    w.u = w.ubase + w.lo;
    w.w = (float*)ctx->ws("mx.w", f);
    w.p = (float*)ctx->ws("mx.p", f);
    w.s = (float*)ctx->ws("mx.s", f);
    w.d = (float*)ctx->ws("mx.d", f);
    w.invd = (float*)ctx->ws("mx.invd", f);
    w.st = (MixedState*)ctx->ws("mx.state", sizeof(MixedState));
    return w;
}

int64_t mixed_blocks(int64_t n) { return std::max<int64_t>(1, (n + kMxTiles * kBlock - 1) / (kMxTiles * kBlock)); }

int64_t mixed_partials(int64_t n) {
    const int64_t nb = std::max<int64_t>(1, nblocks(n));
    return nb + 2 * mixed_blocks(n);
}

void mixed_launch_inner_iter(rsd_ctx* ctx, const rsd_matrix* A, int32_t precond, const MixedWork& w) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = std::max<int64_t>(1, nblocks(n)), nbu = mixed_blocks(n);
    double* P = ctx->partials(mixed_partials(n));
    const float* invd = precond == RSD_PC_JACOBI ? w.invd : nullptr;
    hipStream_t s = ctx->stream;
    if (w.full)
        RSD_LAUNCH_COLS(S.mode(), k_mx_spmv_wu, dim3((unsigned)nb), dim3(kBlock), s, n, nb, S.cols(), A->val32f.p, w.u,
                        w.w, w.st, P);
    else
        RSD_LAUNCH_COLS32(w.lo || w.hi ? S.mode_padded() : S.mode32(), k_mx_spmv_wu, dim3((unsigned)nb),
                          dim3(kBlock), s, n, nb, S.cols32(), A->val32.p, w.u, w.w, w.st, P);
    hipLaunchKernelGGL(k_mx_scalars, dim3(1), dim3(kFold), 0, s, nb, P, nbu, P + nb, w.st);
    hipLaunchKernelGGL(k_mx_update, dim3((unsigned)nbu), dim3(kBlock), 0, s, n, nbu, w.p, w.s, w.d, w.u, w.w, invd,
                       w.st, P + nb);
}

constexpr int64_t kInnerChunk = 8;

// Full mixed solve; A must hold current FP64 values (FP32 copy refreshed here
// unless fresh_f32 is already true).  Returns the final state.
MixedState mixed_run(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                     int32_t precond, double inner_floor, bool fresh_f32, const MixedWork& w, ChunkRunner* runner,
                     int64_t* hint, const double* ax0) {
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = std::max<int64_t>(1, nblocks(n)), nbt = mixed_blocks(n);
    hipStream_t st = ctx->stream;
    if (w.full) {
        if (A->val32f.n != S.slots) A->val32f.alloc(S.slots);
        if (S.slots)
            hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nblocks(S.slots)), dim3(kBlock), 0, st, S.slots, A->val.p,
                               A->val32f.p);
    } else if (!fresh_f32) {
        matrix_to_f32(ctx, A, true);
    }
    double* part = ctx->partials(mixed_partials(n));
    const float* invd = precond == RSD_PC_JACOBI ? w.invd : nullptr;
    if (precond == RSD_PC_JACOBI)
        hipLaunchKernelGGL(k_inv_diag_f32, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, st, n, S.diag.p, A->val.p, w.invd);
    // the padding around u holds zeros (no kernel writes it)
    launch_fill_f32(ctx, w.lo, w.ubase, 0.0f);
    launch_fill_f32(ctx, w.hi, w.u + n, 0.0f);
    MixedState* h = nullptr;
    RSD_HIP(hipHostMalloc((void**)&h, sizeof(MixedState), hipHostMallocDefault));
    std::memset(h, 0, sizeof(MixedState));
    h->in_max = max_iters;
    RSD_HIP(hipMemcpyAsync(w.st, h, sizeof(MixedState), hipMemcpyHostToDevice, st));
    MixedState out{};
    auto chunk = [&] {
        if (runner) runner->run();
        else for (int64_t k = 0; k < kInnerChunk; ++k) mixed_launch_inner_iter(ctx, A, precond, w);
    };
    try {
        // Per cycle: the residual and the outer decision (device), the inner
        // set-up and the chunks the same cycle of the previous solve needed,
        // then one look at the state; every kernel returns at once once the
        // solve is done or the inner solve has stopped.
        const bool known = hint && hint[0] > 0;
        for (int cycle = 0;; ++cycle) {
            // the outer decision ends the loop by cycle kMixedMaxCycles; hint has kMixedMaxCycles + 1 entries
            RSD_CHECK(cycle <= kMixedMaxCycles, "mixed CG: outer loop did not end");
            if (cycle == 0 && ax0)
                hipLaunchKernelGGL(k_mx_residual_ax, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb, b, ax0, w.r64,
                                   part, w.st);
            else
                RSD_LAUNCH_COLS(S.mode(), k_mx_residual, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(),
                                A->val.p, b, x, w.r64, part, w.st);
            hipLaunchKernelGGL(k_mx_outer, dim3(1), dim3(kFold), 0, st, nb, part,
                               S.has_half && !w.full ? A->asym.p : nullptr, w.st, tol, max_iters, inner_floor);
            hipLaunchKernelGGL(k_mx_init, dim3((unsigned)nbt), dim3(kBlock), 0, st, n, nbt, w.r64, w.u, w.p, w.s,
                               w.d, invd, w.st, part + nb);
            const int64_t before = h->total_inner;
            const int64_t hc = hint ? hint[cycle] : 0;
            int64_t launched = 0;
            if (!known) {
                chunk();
                launched = kInnerChunk;
            } else if (hc > 0) {
                // the same cycle of the previous solve's iterations and one
                // more: whole chunks, then single iterations, so the look at
                // the state usually finds the inner solve stopped with at most
                // a few launches left over
                const int64_t want = hc + 1;
                for (; launched + kInnerChunk <= want; launched += kInnerChunk) chunk();
                for (; launched < want; ++launched) mixed_launch_inner_iter(ctx, A, precond, w);
            }
            for (;;) {
                RSD_HIP(hipMemcpyAsync(h, w.st, sizeof(MixedState), hipMemcpyDeviceToHost, st));
                RSD_HIP(hipStreamSynchronize(st));
                if (h->asym)
                    throw Error(RSD_ERR_ARG, "mixed CG: the matrix is not symmetric (a value differs from its mirror by "
                                             "more than 1e-10 of its row's largest entry)");
                if (h->done || h->stop) break;
                RSD_CHECK(launched <= max_iters + 2 * kInnerChunk, "mixed CG: inner solve did not stop");
                chunk();
                launched += kInnerChunk;
            }
            if (hint) hint[cycle] = h->total_inner - before;
            if (h->done) break;
            hipLaunchKernelGGL(k_mx_update_x, dim3((unsigned)nblocks(n)), dim3(kBlock), 0, st, n, x, w.d, w.st);
            RSD_HIP(hipGetLastError());
        }
        out = *h;
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
    return out;
}

}  // namespace rsd

using namespace rsd;

extern "C" {

int rsd_matrix_to_f32(rsd_ctx* ctx, rsd_matrix* A) {
    return guard([&] {
        RSD_CHECK(ctx && A, "null argument");
        RSD_LIVE_MATRIX(A);
        matrix_to_f32(ctx, A, true);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_spmv_f32(rsd_ctx* ctx, rsd_matrix* A, const float* x, float* y) {
    return guard([&] {
        RSD_CHECK(ctx && A && x && y, "null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(A->val32.n == A->pb->sell.slots32(), "rsd_spmv_f32: call rsd_matrix_to_f32 first");
        launch_spmv_f32(ctx, A, x, y, false);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_spmv_f32_padded(rsd_ctx* ctx, rsd_matrix* A, const float* x, float* y) {
    return guard([&] {
        RSD_CHECK(ctx && A && x && y, "null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(A->val32.n == A->pb->sell.slots32(), "rsd_spmv_f32_padded: call rsd_matrix_to_f32 first");
        launch_spmv_f32(ctx, A, x, y, true);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_cg_solve_mixed(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                       int32_t precond, double inner_floor, rsd_report* rep, int64_t* outer_cycles) {
    return guard([&] {
        RSD_CHECK(ctx && A && b && x, "rsd_cg_solve_mixed: null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        MixedWork w = mixed_workspace(ctx, A->pb->sell.n_rows, &A->pb->sell);
        MixedState s = mixed_run(ctx, A, b, x, tol, max_iters, precond,
                                 inner_floor > 0 ? inner_floor : kMixedDefaultInnerFloor, false, w,
                                 nullptr);
        if (rep) {
            rep->converged = s.converged;
            rep->iterations = s.total_inner;
            rep->initial_residual = s.res0;
            rep->final_residual = s.res;
            rep->final_residual_rel = s.res / (s.res0 + 1e-50);
        }
        if (outer_cycles) *outer_cycles = s.cycles;
    });
}

}  // extern "C"
