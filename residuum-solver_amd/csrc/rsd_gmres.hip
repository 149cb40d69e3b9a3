// rsd_gmres.hip -- right-preconditioned restarted GMRES(m) on the device
// (linalg/solver/iterative/gmres/Solver.tpp:5-170).
//
// Modified Gram-Schmidt runs as one fused kernel per basis vector: it
// finishes the previous projection w -= h_{i-1,j} q_{i-1} element by element
// and accumulates q_i . w for the next coefficient, which is exactly the
// reference's dot-then-axpy sequence.  A context may choose classical
// Gram-Schmidt with one reorthogonalisation instead (RSD_ORTHO_CGS2: three
// passes and three folds per step; not bitwise the reference).  The Givens
// rotations run in the block that folds the step's norm, from an LDS copy
// of the column, and the back-substitution in one device thread: scalar
// work in the reference's order.  A whole cycle is launched at once: after the stop flag is set
// the remaining kernels of the cycle return immediately, and the host reads
// the state once per cycle.
#include "rsd_gmres.hpp"
#include "rsd_linalg.hpp"
#include "rsd_mixed.hpp"
#include "rsd_rowdot.hpp"

#include <algorithm>
#include <cmath>

namespace rsd {

// w = A z, z = M q: as the reference, the Jacobi product is a vector,
// formed where q_j is (k_gm_cycle_start, k_gm_newq), so the row product
// gathers one vector with batched loads (row_dot_f64).  Gathering q and
// invd per column and multiplying there gives the same bits but took 83 us
// against 49 us at 1M Q2 rows.
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_gm_apply(int64_t n, int64_t nb, Cols C, const double* __restrict__ val,
                                                     const double* __restrict__ z, double* __restrict__ w,
                                                     const GMState* st) {
    if (st->stop) return;
    const int64_t r = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    if (r >= n) return;
    w[r] = row_dot_f64<MODE>(r, C, val, z);
}

// The same product over the slice-union copy (row_dot_su: the rows of a
// slice share their column offsets; bitwise the sliced product).  At 1M Q2
// rows it took 34.8 us against 49.5 us for the sliced product, whose lanes
// each followed their own offset pattern.
template<typename T>
__global__ __launch_bounds__(kBlock) void k_gm_apply_su(int64_t n, int64_t nb, SuCols C, const T* __restrict__ val,
                                                        const T* __restrict__ z, T* __restrict__ w,
                                                        const GMState* st) {
    if (st->stop) return;
    const int64_t r = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    if ((r >> 6) >= C.n_slices) return;  // whole wavefronts
    const T y = row_dot_su<T>(r, n, C, val, z);
    if (r < n) w[r] = y;
}

// the slice-union copy of the sliced values (FP64 as is, or rounded to
// FP32 as the mixed solver's full copy is): one thread per slice lane
template<typename T>
__global__ __launch_bounds__(kBlock) void k_to_su(int64_t n, SuCols C, const int64_t* __restrict__ slice_ptr,
                                                  const double* __restrict__ val, T* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t s = r >> 6;
    if (s >= C.n_slices) return;
    const int lane = (int)(r & 63);
    const int nu = C.nu[s];
    const int64_t base = C.ptr[s];
    const uint32_t m = r < n ? C.mask[r] : 0u;
    const double* src = val + slice_ptr[s] + lane;
    int j = 0;
    for (int k = 0; k < nu; ++k) {
        T v = T(0);
        if ((m >> k) & 1u) v = (T)src[(int64_t)kSlice * j++];
        out[su_slot(base, k, nu, lane)] = v;
    }
}

// refresh the slice-union copy of A's values (allocated on first use)
template<typename T>
static void launch_to_su(rsd_ctx* ctx, const rsd_matrix* A, DBuf<T>& copy) {
    const Sell& S = A->pb->sell;
    if (copy.n != S.su_slots) copy.alloc(S.su_slots);
    if (S.n_slices)
        hipLaunchKernelGGL(k_to_su<T>, dim3((unsigned)nblocks(S.n_slices * kSlice)), dim3(kBlock), 0, ctx->stream,
                           S.n_rows, S.su_cols(), S.slice_ptr.p, A->val.p, copy.p);
    RSD_HIP(hipGetLastError());
}

// MGS step i of Arnoldi column j: block partials of q_i . w (of w . w for
// the norm step, Qi == nullptr)
__global__ __launch_bounds__(kBlock) void k_gm_mgs(int64_t n, int64_t nb, double* __restrict__ w,
                                                   const double* __restrict__ Qprev, const double* __restrict__ Qi,
                                                   int32_t i, int32_t j, int32_t m, const double* H, const GMState* st,
                                                   double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t k = lb * kBlock + threadIdx.x;
    const double hprev = i > 0 ? H[(int64_t)j * (m + 1) + (i - 1)] : 0.0;
    double v[1] = {0.0};
    if (k < n) {
        double wk = w[k];
        if (i > 0) {
            wk = wk + -hprev * Qprev[k];
            w[k] = wk;
        }
        v[0] = Qi ? Qi[k] * wk : wk * wk;
    }
    block_partials<1>(v, lb, nb, partials);
}

constexpr int kGivensMax = 512;  // the largest Krylov dimension the entry points accept

// Rotations of column j by the previous j rotations, the new rotation and
// the residual estimate (Solver.tpp:78-122), run by the whole block of a
// norm fold once hnext is known (in thread 0): the threads stage column j
// (plus cadd when given: CGS2's second coefficients), cs and sn in LDS
// together, then thread 0 runs the sequential recurrence on the staged
// copy.  A one-thread kernel made one dependent global round trip per
// rotation; here the loads are one parallel batch.  Same operations in
// the same order.
__device__ __forceinline__ void givens_column(int32_t j, int32_t m, double* H, const double* cadd, double* cs,
                                              double* sn, double* g, GMState* st, double hnext, double tol,
                                              int32_t abs_tol, int64_t max_iters) {
    __shared__ double col[kGivensMax], c[kGivensMax], s[kGivensMax];
    double* hj = H + (int64_t)j * (m + 1);
    for (int i = threadIdx.x; i <= j; i += blockDim.x) col[i] = cadd ? hj[i] + cadd[i] : hj[i];
    for (int i = threadIdx.x; i < j; i += blockDim.x) {
        c[i] = cs[i];
        s[i] = sn[i];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    st->total += 1;
    const int32_t breakdown = hnext <= 1e-14;
    for (int i = 0; i < j; ++i) {
        const double h1 = col[i], h2 = col[i + 1];
        col[i] = c[i] * h1 + s[i] * h2;
        col[i + 1] = -s[i] * h1 + c[i] * h2;
    }
    const double hjj = col[j];
    const double den = sqrt(hjj * hjj + hnext * hnext);
    double cj = 1.0, sj = 0.0;
    if (!(den <= 1e-14)) {
        cj = hjj / den;
        sj = hnext / den;
    }
    cs[j] = cj;
    sn[j] = sj;
    col[j] = cj * hjj + sj * hnext;
    for (int i = 0; i <= j; ++i) hj[i] = col[i];
    const double gj = g[j];
    g[j] = cj * gj;
    const double gn = -sj * gj;
    g[j + 1] = gn;
    const double est = fabs(gn);
    st->est = est;
    st->mloc = j + 1;
    st->breakdown = breakdown;
    if ((abs_tol ? est : est / (st->res0 + 1e-50)) < tol || breakdown || st->total >= max_iters) st->stop = 1;
}

__global__ __launch_bounds__(kFold) void k_gm_mgs_fold(int64_t nb, const double* __restrict__ partials, int32_t i,
                                                        int32_t j, int32_t m, double* H, GMState* st) {
    if (st->stop) return;
    double out[1];
    fold_partials<1>(nb, partials, out);
    if (threadIdx.x == 0) H[(int64_t)j * (m + 1) + i] = out[0];
}

// the norm step's fold: h_{j+1,j} = ||w||, then column j's rotations
__global__ __launch_bounds__(kFold) void k_gm_mgs_norm(int64_t nb, const double* __restrict__ partials, int32_t j,
                                                        int32_t m, double* H, double* cs, double* sn, double* g,
                                                        GMState* st, double tol, int32_t abs_tol, int64_t max_iters) {
    if (st->stop) return;
    double out[1];
    fold_partials<1>(nb, partials, out);
    const double hnext = sqrt(out[0]);
    if (threadIdx.x == 0) st->hnext = hnext;
    givens_column(j, m, H, nullptr, cs, sn, g, st, hnext, tol, abs_tol, max_iters);
}

// ------------------------------------------------------------------------
// Classical Gram-Schmidt with one reorthogonalisation (CGS2, the context's
// RSD_ORTHO_CGS2; not the reference's modified Gram-Schmidt, so results
// differ from it by rounding).  Per Arnoldi step, after w = A M q_j:
//   h  = Q^T w                 (k_gm_dots: one pass over w and the basis)
//   w -= Q h, c = Q^T w        (k_gm_upd_dots: one pass)
//   w -= Q c, ||w||^2          (k_gm_upd_norm: one pass)
//   H[:, j] = h + c, h_{j+1,j} = ||w||
// -- three block-sum folds per step instead of j + 2.  Each dot's block sums
// use the fixed tree of every other reduction here.
// ------------------------------------------------------------------------
constexpr int kCgsMax = 32;  // basis vectors per pass of the multi-vector kernels

// DPP lane moves of a double (bound_ctrl: a lane with no source reads 0;
// lanes outside the row mask read 0 as well)
template<int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
    const int2 s = __builtin_bit_cast(int2, v);
    int2 d;
    d.x = __builtin_amdgcn_update_dpp(0, s.x, CTRL, ROWS, 0xf, true);
    d.y = __builtin_amdgcn_update_dpp(0, s.y, CTRL, ROWS, 0xf, true);
    return __builtin_bit_cast(double, d);
}

// Wavefront sum in a fixed tree of VALU lane moves: row_shr 1, 2, 4, 8
// within each 16-lane row (lane 15 of a row then holds the row's sum),
// row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3; lane 63
// ends with the sum.  A many-value block sum through __shfl_down (two
// ds_bpermute per step and value) cost the multi-dot kernel a third of its
// time.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v = v + dpp_f64<0x111, 0xf>(v);
    v = v + dpp_f64<0x112, 0xf>(v);
    v = v + dpp_f64<0x114, 0xf>(v);
    v = v + dpp_f64<0x118, 0xf>(v);
    v = v + dpp_f64<0x142, 0xa>(v);
    v = v + dpp_f64<0x143, 0xc>(v);
    return v;
}

// block sums of cnt (<= kCgsMax, which the launches keep) values per thread, value c's sum of block lb
// stored at partials[c * nb + lb]: each value's four wavefront sums
// (wave_sum_dpp), then thread c adds them left to right
__device__ __forceinline__ void block_partials_many(double (&v)[kCgsMax], int cnt, int64_t lb, int64_t nb,
                                                    double* partials) {
    __shared__ double sm[kCgsMax * 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < kCgsMax; ++c) {
        if (c < cnt) {  // cnt is uniform: whole wavefronts take the lane moves
            const double t = wave_sum_dpp(v[c]);
            if (lane == 63) sm[c * 4 + wave] = t;
        }
    }
    __syncthreads();
    const int c = threadIdx.x;
    if (c < cnt) {
        double r = sm[c * 4 + 0];
        r += sm[c * 4 + 1];
        r += sm[c * 4 + 2];
        r += sm[c * 4 + 3];
        partials[(int64_t)c * nb + lb] = r;
    }
}

// The multi-dot kernels give each thread kCgsTiles rows, kBlock apart (its
// products added in row order), so a block's cnt-value reduction is paid once
// per kCgsTiles * kBlock rows: with one row per thread it took more time than
// the block's loads.  More tiles mean fewer blocks and fewer waves in flight:
// on vectors that stay in the last-level cache 16 tiles halved these kernels'
// time, but in Newton's GMRES(30) at 1M rows, where the basis streams from
// HBM, a step took 161 / 161 / 179 us (FP64) and 107 / 113 / 132 us (mixed)
// with 4 / 8 / 16 tiles.
#ifndef RSD_CGS_TILES
#define RSD_CGS_TILES 4
#endif
constexpr int kCgsTiles = RSD_CGS_TILES;
inline int64_t cgs_blocks(int64_t n) { return std::max<int64_t>(1, (n + kCgsTiles * kBlock - 1) / (kCgsTiles * kBlock)); }

// The basis type's arithmetic: FP64 basis vectors (the FP64 solver) update
// w with separately rounded products and sums; FP32 ones (the mixed solver's
// inner GMRES) with FP32 fused multiply-adds, as its MGS kernel does.  Dots
// sum the exact FP64 products of both.
__device__ __forceinline__ double cgs_sub(double wk, double h, double q) { return wk + -h * q; }
__device__ __forceinline__ float cgs_sub(float wk, double h, float q) { return __builtin_fmaf(-(float)h, q, wk); }

// q[c] = Q[i0 + c][k] for c < cnt, 0 past it: every load is issued before
// any of them is used (a load and its use under one guard would wait for
// each load in turn)
template<typename T>
__device__ __forceinline__ void cgs_load(T (&q)[kCgsMax], const T* const* Q, int32_t i0, int32_t cnt, int64_t k) {
#pragma unroll
    for (int c = 0; c < kCgsMax; ++c) {
        q[c] = T(0);
        if (c < cnt) q[c] = Q[i0 + c][k];
    }
}

// block sums of q_{i0+c} . w, c < cnt
template<typename T>
__global__ __launch_bounds__(kBlock) void k_gm_dots(int64_t n, int64_t nb, const T* __restrict__ w,
                                                    const T* const* __restrict__ Q, int32_t i0, int32_t cnt,
                                                    const GMState* st, double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    double v[kCgsMax];
#pragma unroll
    for (int c = 0; c < kCgsMax; ++c) v[c] = 0.0;
    // FP64: tiles unrolled; FP32: one tile at a time (unrolled, the compiler
    // hoisted every tile's loads and spilled 1.8 KB per thread)
    constexpr int kTileUnroll = sizeof(T) == 8 ? kCgsTiles : 1;
#pragma unroll kTileUnroll
    for (int t = 0; t < kCgsTiles; ++t) {
        const int64_t k = (lb * kCgsTiles + t) * kBlock + threadIdx.x;
        if (k < n) {
            const double wk = (double)w[k];
            T q[kCgsMax];
            cgs_load(q, Q, i0, cnt, k);
#pragma unroll
            for (int c = 0; c < kCgsMax; ++c)
                if (c < cnt) v[c] += (double)q[c] * wk;
        }
    }
    block_partials_many(v, cnt, lb, nb, partials + (int64_t)i0 * nb);
}

// w -= sum_i h_i q_i (i = 0..cnt-1 in order), then, with dots, the block sums
// of q_i . w for the new w, or, without, of w . w
template<typename T>
__device__ __forceinline__ T cgs_subtract(int64_t k, const T* w, const T* const* Q, const double* h, int32_t cnt) {
    T wk = w[k];
#pragma unroll 8
    for (int i = 0; i < cnt; ++i) wk = cgs_sub(wk, h[i], Q[i][k]);
    return wk;
}

// (the dots of q_0 .. q_{cd-1}, cd <= kCgsMax; k_gm_dots forms the rest)
template<typename T>
__global__ __launch_bounds__(kBlock) void k_gm_upd_dots(int64_t n, int64_t nb, T* __restrict__ w,
                                                        const T* const* __restrict__ Q, const double* h,
                                                        int32_t cnt, int32_t cd, const GMState* st,
                                                        double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    double v[kCgsMax];
#pragma unroll
    for (int c = 0; c < kCgsMax; ++c) v[c] = 0.0;
#pragma unroll
    for (int t = 0; t < kCgsTiles; ++t) {
        const int64_t k = (lb * kCgsTiles + t) * kBlock + threadIdx.x;
        if (k < n) {
            T q[kCgsMax];
            T wk = w[k];
            // w -= sum_i h_i q_i in order, kCgsMax basis entries at a time; the
            // dots then use the first chunk (loaded again if a later one replaced it)
            for (int32_t i0 = 0; i0 < cnt; i0 += kCgsMax) {
                const int32_t cc = cnt - i0 < kCgsMax ? cnt - i0 : kCgsMax;
                cgs_load(q, Q, i0, cc, k);
#pragma unroll
                for (int c = 0; c < kCgsMax; ++c)
                    if (c < cc) wk = cgs_sub(wk, h[i0 + c], q[c]);
            }
            w[k] = wk;
            if (cnt > kCgsMax) cgs_load(q, Q, 0, cd, k);
#pragma unroll
            for (int c = 0; c < kCgsMax; ++c)
                if (c < cd) v[c] += (double)q[c] * (double)wk;
        }
    }
    block_partials_many(v, cd, lb, nb, partials);
}

template<typename T>
__global__ __launch_bounds__(kBlock) void k_gm_upd_norm(int64_t n, int64_t nb, T* __restrict__ w,
                                                        const T* const* __restrict__ Q, const double* h,
                                                        int32_t cnt, const GMState* st, double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t k = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (k < n) {
        const T wk = cgs_subtract(k, w, Q, h, cnt);
        w[k] = wk;
        v[0] = (double)wk * (double)wk;
    }
    block_partials<1>(v, lb, nb, partials);
}

// one block per array: out[c] = the fold of partials[c * nb ..]
__global__ __launch_bounds__(kFold) void k_gm_fold_many(int64_t nb, const double* __restrict__ partials,
                                                        double* out, const GMState* st) {
    if (st->stop) return;
    double o[1];
    fold_partials<1>(nb, partials + (int64_t)blockIdx.x * nb, o);
    if (threadIdx.x == 0) out[blockIdx.x] = o[0];
}

// h_{j+1,j} = ||w|| and H[:, j] = h + c, then column j's rotations (one block)
__global__ __launch_bounds__(kFold) void k_gm_norm_fold(int64_t nb, const double* __restrict__ partials, int32_t j,
                                                        int32_t m, const double* __restrict__ c, double* H, double* cs,
                                                        double* sn, double* g, GMState* st, double tol,
                                                        int32_t abs_tol, int64_t max_iters) {
    if (st->stop) return;
    double o[1];
    fold_partials<1>(nb, partials, o);
    const double hnext = sqrt(o[0]);
    if (threadIdx.x == 0) st->hnext = hnext;
    givens_column(j, m, H, c, cs, sn, g, st, hnext, tol, abs_tol, max_iters);
}

// q_{j+1} = w * (1 / hnext) unless the column broke down
__global__ void k_gm_newq(int64_t n, const double* __restrict__ w, double* __restrict__ q,
                          const double* __restrict__ invd, double* __restrict__ z, const GMState* st) {
    if (st->stop || st->hnext <= 1e-14) return;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        const double s = 1.0 / st->hnext;
        const double qk = w[k] * s;
        q[k] = qk;
        if (invd) z[k] = qk * invd[k];  // z = M q_{j+1} for the next product
    }
}

// back-substitution H y = g (Solver.tpp:126-133)
__global__ void k_gm_backsub(int32_t m, const double* H, const double* g, double* y, const GMState* st) {
    const int mloc = st->mloc;
    for (int ii = 0; ii < mloc; ++ii) {
        const int i = mloc - 1 - ii;
        double s = g[i];
        for (int k = i + 1; k < mloc; ++k) s -= H[(int64_t)k * (m + 1) + i] * y[k];
        y[i] = s / H[(int64_t)i * (m + 1) + i];
    }
}

// x += M (sum_i y_i q_i)  (Solver.tpp:135-139)
__global__ void k_gm_update_x(int64_t n, const double* const* __restrict__ Q, const double* __restrict__ y,
                              const double* __restrict__ invd, double* __restrict__ x, const GMState* st) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    double z = 0.0;
    for (int i = 0; i < st->mloc; ++i) z += y[i] * Q[i][k];
    const double wk = invd ? z * invd[k] : z;
    x[k] += 1.0 * wk;
}

// q0 = b - A x, beta = ||q0|| (Solver.tpp:22-27, 141-147)
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_gm_residual(int64_t n, int64_t nb, Cols C, const double* __restrict__ val,
                                                        const double* __restrict__ b, const double* __restrict__ x,
                                                        double* __restrict__ q0, double* partials) {
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t k = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (k < n) {
        const double qk = b[k] + -1.0 * row_dot_f64<MODE>(k, C, val, x);
        q0[k] = qk;
        v[0] = qk * qk;
    }
    block_partials<1>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kFold) void k_gm_residual_fold(int64_t nb, const double* __restrict__ partials,
                                                             GMState* st, int32_t first) {
    double out[1];
    fold_partials<1>(nb, partials, out);
    if (threadIdx.x == 0) {
        const double beta = sqrt(out[0]);
        st->beta = beta;
        if (first) st->res0 = beta;
    }
}

// start of a cycle: q0 *= 1/beta (and z = M q0), g = (beta, 0, ...)
__global__ void k_gm_cycle_start(int64_t n, double* __restrict__ q0, const double* __restrict__ invd,
                                 double* __restrict__ z, double* g, int32_t m, GMState* st) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double beta = st->beta;
    if (k < n) {
        const double qk = q0[k] * (1.0 / beta);
        q0[k] = qk;
        if (invd) z[k] = qk * invd[k];
    }
    if (k == 0) {
        g[0] = beta;
        for (int i = 1; i <= m; ++i) g[i] = 0.0;
    }
}

__global__ void k_gm_reset_flags(GMState* st) {
    st->stop = 0;
    st->breakdown = 0;
    st->mloc = 0;
}

GMRESResult gmres_run(rsd_ctx* ctx, const rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                      int32_t m, int32_t precond, bool abs_tol) {
    auto crit = [&](double res, double res0) { return abs_tol ? res : res / (res0 + 1e-50); };
    RSD_CHECK(m >= 1 && m <= kGivensMax, "GMRES: Krylov dimension outside [1, 512]");  // the rotations' LDS copy
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = std::max<int64_t>(1, nblocks(n)), nbe = nblocks(std::max<int64_t>(n, 1));
    hipStream_t st = ctx->stream;
    const size_t vb = sizeof(double) * (size_t)std::max<int64_t>(n, 1);
    std::vector<double*> hq(m + 1);
    for (int i = 0; i <= m; ++i) hq[i] = (double*)ctx->ws("gm.Q" + std::to_string(i), vb);
    double* w = (double*)ctx->ws("gm.w", vb);
    double* invd = precond == RSD_PC_JACOBI ? (double*)ctx->ws("gm.invd", vb) : nullptr;
    double* z = invd ? (double*)ctx->ws("gm.z", vb) : nullptr;  // M q_j
    const bool su = S.has_su;  // the products read the slice-union copy
    double* H = (double*)ctx->ws("gm.H", sizeof(double) * (size_t)(m + 1) * m);
    double* cs = (double*)ctx->ws("gm.cs", sizeof(double) * (size_t)(m + 1));
    double* sn = (double*)ctx->ws("gm.sn", sizeof(double) * (size_t)(m + 1));
    double* g = (double*)ctx->ws("gm.g", sizeof(double) * (size_t)(m + 1));
    double* y = (double*)ctx->ws("gm.y", sizeof(double) * (size_t)(m + 1));
    double** dQ = (double**)ctx->ws("gm.Qptr", sizeof(double*) * (size_t)(m + 1));
    GMState* dst = (GMState*)ctx->ws("gm.state", sizeof(GMState));
    const bool cgs2 = ctx->gmres_orth == RSD_ORTHO_CGS2;
    double* hc = cgs2 ? (double*)ctx->ws("gm.hc", sizeof(double) * (size_t)(m + 1)) : nullptr;
    // MGS: one dot at a time; CGS2: up to m dots at a time
    const int64_t nbc = cgs_blocks(n);  // blocks of the multi-dot kernels
    double* part = ctx->partials(cgs2 ? std::max<int64_t>((int64_t)m * nbc, nb) : nb);
    RSD_HIP(hipMemcpyAsync(dQ, hq.data(), sizeof(double*) * (size_t)(m + 1), hipMemcpyHostToDevice, st));
    RSD_HIP(hipMemsetAsync(dst, 0, sizeof(GMState), st));
    if (invd) launch_inv_diag(ctx, A, invd);
    if (su) launch_to_su(ctx, A, A->val_su);
    GMState* h = nullptr;
    RSD_HIP(hipHostMalloc((void**)&h, sizeof(GMState), hipHostMallocDefault));
    auto read = [&] {
        RSD_HIP(hipMemcpyAsync(h, dst, sizeof(GMState), hipMemcpyDeviceToHost, st));
        RSD_HIP(hipStreamSynchronize(st));
    };
    GMRESResult res{};
    try {
        RSD_LAUNCH_COLS(S.mode(), k_gm_residual, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(), A->val.p, b, x,
                        hq[0], part);
        hipLaunchKernelGGL(k_gm_residual_fold, dim3(1), dim3(kFold), 0, st, nb, part, dst, 1);
        read();
        const double res0 = h->res0;
        res.initial_residual = res0;
        if (crit(res0, res0) < tol) {
            res.converged = 1;
            res.final_residual = res0;
            res.final_residual_rel = 1.0;
            res.iterations = 0;
        } else {
            int64_t total = 0;
            // one restart cycle: every launch returns at once after the
            // state's stop flag is set, and its arguments are the same in
            // every cycle, so from the second cycle on it is replayed from a
            // graph (about 2 m + m^2 / 2 launches per cycle)
            auto cycle = [&] {
                hipLaunchKernelGGL(k_gm_reset_flags, dim3(1), dim3(1), 0, st, dst);
                hipLaunchKernelGGL(k_gm_cycle_start, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, hq[0], invd, z, g, m,
                                   dst);
                for (int j = 0; j < m; ++j) {
                    if (su)
                        hipLaunchKernelGGL(k_gm_apply_su<double>, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb,
                                           S.su_cols(), (const double*)A->val_su.p, invd ? z : hq[j], w, dst);
                    else
                        RSD_LAUNCH_COLS(S.mode(), k_gm_apply, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(),
                                        A->val.p, invd ? z : hq[j], w, dst);
                    if (cgs2) {
                        // h = Q^T w (passes of kCgsMax vectors), w -= Q h with c = Q^T w,
                        // w -= Q c with ||w||^2, H[:, j] = h + c
                        const int32_t cnt = j + 1;
                        double* hcol = H + (int64_t)j * (m + 1);
                        for (int32_t i0 = 0; i0 < cnt; i0 += kCgsMax)
                            hipLaunchKernelGGL(k_gm_dots<double>, dim3((unsigned)nbc), dim3(kBlock), 0, st, n, nbc, w,
                                               (const double* const*)dQ, i0, std::min(kCgsMax, cnt - i0), dst, part);
                        hipLaunchKernelGGL(k_gm_fold_many, dim3((unsigned)cnt), dim3(kFold), 0, st, nbc, part, hcol,
                                           dst);
                        hipLaunchKernelGGL(k_gm_upd_dots<double>, dim3((unsigned)nbc), dim3(kBlock), 0, st, n, nbc, w,
                                           (const double* const*)dQ, hcol, cnt, std::min(kCgsMax, cnt), dst, part);
                        for (int32_t i0 = kCgsMax; i0 < cnt; i0 += kCgsMax)
                            hipLaunchKernelGGL(k_gm_dots<double>, dim3((unsigned)nbc), dim3(kBlock), 0, st, n, nbc, w,
                                               (const double* const*)dQ, i0, std::min(kCgsMax, cnt - i0), dst, part);
                        hipLaunchKernelGGL(k_gm_fold_many, dim3((unsigned)cnt), dim3(kFold), 0, st, nbc, part, hc,
                                           dst);
                        hipLaunchKernelGGL(k_gm_upd_norm<double>, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb, w,
                                           (const double* const*)dQ, hc, cnt, dst, part);
                        hipLaunchKernelGGL(k_gm_norm_fold, dim3(1), dim3(kFold), 0, st, nb, part, j, m, hc, H, cs, sn,
                                           g, dst, tol, (int32_t)abs_tol, max_iters);
                    } else {
                        for (int i = 0; i <= j + 1; ++i) {
                            const bool norm = i == j + 1;
                            hipLaunchKernelGGL(k_gm_mgs, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb, w,
                                               i > 0 ? hq[i - 1] : nullptr, norm ? nullptr : hq[i], i, j, m, H, dst,
                                               part);
                            if (norm)
                                hipLaunchKernelGGL(k_gm_mgs_norm, dim3(1), dim3(kFold), 0, st, nb, part, j, m, H, cs,
                                                   sn, g, dst, tol, (int32_t)abs_tol, max_iters);
                            else
                                hipLaunchKernelGGL(k_gm_mgs_fold, dim3(1), dim3(kFold), 0, st, nb, part, i, j, m, H,
                                                   dst);
                        }
                    }
                    // the norm fold ran the rotations: q_{j+1} only when the cycle goes on
                    hipLaunchKernelGGL(k_gm_newq, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, w, hq[j + 1], invd, z,
                                       dst);
                    RSD_HIP(hipGetLastError());
                }
                hipLaunchKernelGGL(k_gm_backsub, dim3(1), dim3(1), 0, st, m, H, g, y, dst);
                hipLaunchKernelGGL(k_gm_update_x, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, (const double* const*)dQ,
                                   y, invd, x, dst);
                RSD_LAUNCH_COLS(S.mode(), k_gm_residual, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(),
                                A->val.p, b, x, hq[0], part);
                hipLaunchKernelGGL(k_gm_residual_fold, dim3(1), dim3(kFold), 0, st, nb, part, dst, 0);
                RSD_HIP(hipGetLastError());
            };
            Graph cyc;
            for (int64_t c = 0;; ++c) {
                if (c == 0) {
                    cycle();
                } else {
                    if (!cyc.e) capture(st, cyc, cycle);
                    RSD_HIP(hipGraphLaunch(cyc.e, st));
                }
                read();
                total = h->total;
                const double beta = h->beta;
                const bool conv = crit(beta, res0) < tol;
                if (conv || h->breakdown || total >= max_iters) {
                    res.converged = conv;
                    res.final_residual = beta;
                    res.final_residual_rel = beta / (res0 + 1e-50);
                    res.iterations = total;
                    break;
                }
            }
        }
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
    return res;
}

// ------------------------------------------------------------------------
// Mixed precision GMRES: iterative refinement around an FP32 inner solve.
// Outer (FP64): r = b - A x with the FP64 values, stop on ||r|| / (||r_0||
// + 1e-50) < tol, x += ||r|| d.  Inner: right-preconditioned GMRES(m) on
// A d = r / ||r|| with the FP32 copy of the values (any matrix: the full
// sliced ELL layout, not the symmetric half storage), FP32 basis vectors,
// products and vector updates; the Gram-Schmidt and norm dots sum exact
// FP64 products; Hessenberg, Givens and back-substitution in FP64 (the
// kernels of the FP64 solver).  The inner tolerance is the mixed CG's:
// max(floor, 0.5 tol ||r_0|| / ||r||), capped at 0.5.
// ------------------------------------------------------------------------
__global__ void k_gm32_convert(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_gm32_inv_diag(int64_t n, const int32_t* __restrict__ diag, const double* __restrict__ val,
                                float* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t s = diag[i];
    const double d = s >= 0 ? val[s] : 0.0;
    inv[i] = (float)((fabs(d) > 1e-300) ? 1.0 / d : 1.0);
}

// w = A z in FP32 (fused multiply-adds in CSR order), z = M q formed where
// q_j is, as in the FP64 solver
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_gm32_apply(int64_t n, int64_t nb, Cols C, const float* __restrict__ val,
                                                       const float* __restrict__ z, float* __restrict__ w,
                                                       const GMState* st) {
    if (st->stop) return;
    const int64_t r = xcd_block(blockIdx.x, nb) * kBlock + threadIdx.x;
    if (r >= n) return;
    w[r] = row_dot_f32<MODE>(r, C, val, z);
}

// inner residual q0 = r64 / ||r|| - A32 d (d = 0 on the first inner cycle:
// no product), block sums of ||q0||^2
template<int MODE>
__global__ __launch_bounds__(kBlock) void k_gm32_inres(int64_t n, int64_t nb, Cols C, const float* __restrict__ val,
                                                       const double* __restrict__ r64, double sc,
                                                       const float* __restrict__ d, int32_t first,
                                                       float* __restrict__ q0, double* partials) {
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t k = lb * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (k < n) {
        float qk = (float)(r64[k] * sc);
        if (!first) qk = qk - row_dot_f32<MODE>(k, C, val, d);
        q0[k] = qk;
        v[0] = (double)qk * (double)qk;
    }
    block_partials<1>(v, lb, nb, partials);
}

__global__ __launch_bounds__(kBlock) void k_gm32_mgs(int64_t n, int64_t nb, float* __restrict__ w,
                                                     const float* __restrict__ Qprev, const float* __restrict__ Qi,
                                                     int32_t i, int32_t j, int32_t m, const double* H,
                                                     const GMState* st, double* partials) {
    if (st->stop) return;
    const int64_t lb = xcd_block(blockIdx.x, nb);
    const int64_t k = lb * kBlock + threadIdx.x;
    const float hprev = i > 0 ? (float)H[(int64_t)j * (m + 1) + (i - 1)] : 0.0f;
    double v[1] = {0.0};
    if (k < n) {
        float wk = w[k];
        if (i > 0) {
            wk = __builtin_fmaf(-hprev, Qprev[k], wk);
            w[k] = wk;
        }
        v[0] = Qi ? (double)Qi[k] * (double)wk : (double)wk * (double)wk;
    }
    block_partials<1>(v, lb, nb, partials);
}

__global__ void k_gm32_newq(int64_t n, const float* __restrict__ w, float* __restrict__ q,
                            const float* __restrict__ invd, float* __restrict__ z, const GMState* st) {
    if (st->stop || st->hnext <= 1e-14) return;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        const float qk = w[k] * (float)(1.0 / st->hnext);
        q[k] = qk;
        if (invd) z[k] = qk * invd[k];
    }
}

__global__ void k_gm32_cycle_start(int64_t n, float* __restrict__ q0, const float* __restrict__ invd,
                                   float* __restrict__ z, double* g, int32_t m, GMState* st) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double beta = st->beta;
    if (k < n) {
        const float qk = q0[k] * (float)(1.0 / beta);
        q0[k] = qk;
        if (invd) z[k] = qk * invd[k];
    }
    if (k == 0) {
        g[0] = beta;
        for (int i = 1; i <= m; ++i) g[i] = 0.0;
    }
}

// d += M (sum_i y_i q_i), the sum in FP64
__global__ void k_gm32_update_d(int64_t n, const float* const* __restrict__ Q, const double* __restrict__ y,
                                const float* __restrict__ invd, float* __restrict__ d, const GMState* st) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    double z = 0.0;
    for (int i = 0; i < st->mloc; ++i) z += y[i] * (double)Q[i][k];
    if (invd) z *= (double)invd[k];
    d[k] = (float)((double)d[k] + z);
}

__global__ void k_gm32_update_x(int64_t n, double* __restrict__ x, const float* __restrict__ d, double res) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) x[k] += res * (double)d[k];
}

MixedGMResult gmres_mixed_run(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol,
                              int64_t max_iters, int32_t m, int32_t precond, double inner_floor) {
    RSD_CHECK(m >= 1 && m <= kGivensMax, "GMRES: Krylov dimension outside [1, 512]");  // the rotations' LDS copy
    const Sell& S = A->pb->sell;
    const int64_t n = S.n_rows, nb = std::max<int64_t>(1, nblocks(n)), nbe = nblocks(std::max<int64_t>(n, 1));
    hipStream_t st = ctx->stream;
    const size_t fb = sizeof(float) * (size_t)std::max<int64_t>(n, 1);
    std::vector<float*> hq(m + 1);
    for (int i = 0; i <= m; ++i) hq[i] = (float*)ctx->ws("gm32.Q" + std::to_string(i), fb);
    float* w = (float*)ctx->ws("gm32.w", fb);
    float* d = (float*)ctx->ws("gm32.d", fb);
    float* invd = precond == RSD_PC_JACOBI ? (float*)ctx->ws("gm32.invd", fb) : nullptr;
    float* z = invd ? (float*)ctx->ws("gm32.z", fb) : nullptr;  // M q_j
    const bool su = S.has_su;  // the products read the FP32 slice-union copy
    double* r64 = (double*)ctx->ws("gm32.r64", sizeof(double) * (size_t)std::max<int64_t>(n, 1));
    double* H = (double*)ctx->ws("gm.H", sizeof(double) * (size_t)(m + 1) * m);
    double* cs = (double*)ctx->ws("gm.cs", sizeof(double) * (size_t)(m + 1));
    double* sn = (double*)ctx->ws("gm.sn", sizeof(double) * (size_t)(m + 1));
    double* g = (double*)ctx->ws("gm.g", sizeof(double) * (size_t)(m + 1));
    double* y = (double*)ctx->ws("gm.y", sizeof(double) * (size_t)(m + 1));
    float** dQ = (float**)ctx->ws("gm32.Qptr", sizeof(float*) * (size_t)(m + 1));
    GMState* so = (GMState*)ctx->ws("gm32.outer", sizeof(GMState));
    GMState* si = (GMState*)ctx->ws("gm32.inner", sizeof(GMState));
    const bool cgs2 = ctx->gmres_orth == RSD_ORTHO_CGS2;
    double* hc = cgs2 ? (double*)ctx->ws("gm.hc", sizeof(double) * (size_t)(m + 1)) : nullptr;
    const int64_t nbc = cgs_blocks(n);
    double* part = ctx->partials(cgs2 ? std::max<int64_t>((int64_t)m * nbc, nb) : nb);
    if (A->val32f.n != S.slots) A->val32f.alloc(S.slots);
    if (S.slots)
        hipLaunchKernelGGL(k_gm32_convert, dim3((unsigned)nblocks(S.slots)), dim3(kBlock), 0, st, S.slots, A->val.p,
                           A->val32f.p);
    if (su) launch_to_su(ctx, A, A->val_su32);
    if (invd)
        hipLaunchKernelGGL(k_gm32_inv_diag, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, S.diag.p, A->val.p, invd);
    RSD_HIP(hipMemcpyAsync(dQ, hq.data(), sizeof(float*) * (size_t)(m + 1), hipMemcpyHostToDevice, st));
    RSD_HIP(hipMemsetAsync(so, 0, sizeof(GMState), st));
    GMState* h = nullptr;
    RSD_HIP(hipHostMalloc((void**)&h, sizeof(GMState), hipHostMallocDefault));
    auto read = [&](const GMState* s) {
        RSD_HIP(hipMemcpyAsync(h, s, sizeof(GMState), hipMemcpyDeviceToHost, st));
        RSD_HIP(hipStreamSynchronize(st));
    };
    auto outer_residual = [&](int32_t first) {
        RSD_LAUNCH_COLS(S.mode(), k_gm_residual, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(), A->val.p, b, x,
                        r64, part);
        hipLaunchKernelGGL(k_gm_residual_fold, dim3(1), dim3(kFold), 0, st, nb, part, so, first);
        read(so);
        return h->beta;
    };
    MixedGMResult out{};
    try {
        const double res0 = outer_residual(1);
        double res = res0;
        out.initial_residual = res0;
        int64_t total = 0, cycles = 0;
        for (;;) {
            const double rel = res / (res0 + 1e-50);
            if (rel < tol) { out.converged = 1; break; }
            if (cycles >= kMixedMaxCycles || total >= max_iters || !(rel == rel)) break;
            double in_tol = 0.5 * tol * (res0 + 1e-50) / res;
            in_tol = std::min(0.5, std::max(inner_floor, in_tol));
            const int64_t cap = std::max<int64_t>(1, max_iters - total);
            launch_fill_f32(ctx, n, d, 0.0f);
            RSD_HIP(hipMemsetAsync(si, 0, sizeof(GMState), st));
            // one inner restart cycle; its arguments (in_tol, cap) are fixed
            // for this refinement, so from the second restart on it is
            // replayed from a graph
            auto cycle = [&] {
                hipLaunchKernelGGL(k_gm_reset_flags, dim3(1), dim3(1), 0, st, si);
                hipLaunchKernelGGL(k_gm32_cycle_start, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, hq[0], invd, z, g, m,
                                   si);
                for (int j = 0; j < m; ++j) {
                    if (su)
                        hipLaunchKernelGGL(k_gm_apply_su<float>, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb,
                                           S.su_cols(), (const float*)A->val_su32.p, invd ? z : hq[j], w, si);
                    else
                        RSD_LAUNCH_COLS(S.mode(), k_gm32_apply, dim3((unsigned)nb), dim3(kBlock), st, n, nb,
                                        S.cols(), A->val32f.p, invd ? z : hq[j], w, si);
                    if (cgs2) {  // the FP64 solver's three passes on the FP32 basis
                        const int32_t cnt = j + 1;
                        const float* const* Qf = (const float* const*)dQ;
                        double* hcol = H + (int64_t)j * (m + 1);
                        for (int32_t i0 = 0; i0 < cnt; i0 += kCgsMax)
                            hipLaunchKernelGGL(k_gm_dots<float>, dim3((unsigned)nbc), dim3(kBlock), 0, st, n, nbc, w, Qf,
                                               i0, std::min(kCgsMax, cnt - i0), si, part);
                        hipLaunchKernelGGL(k_gm_fold_many, dim3((unsigned)cnt), dim3(kFold), 0, st, nbc, part, hcol,
                                           si);
                        hipLaunchKernelGGL(k_gm_upd_dots<float>, dim3((unsigned)nbc), dim3(kBlock), 0, st, n, nbc, w,
                                           Qf, hcol, cnt, std::min(kCgsMax, cnt), si, part);
                        for (int32_t i0 = kCgsMax; i0 < cnt; i0 += kCgsMax)
                            hipLaunchKernelGGL(k_gm_dots<float>, dim3((unsigned)nbc), dim3(kBlock), 0, st, n, nbc, w, Qf,
                                               i0, std::min(kCgsMax, cnt - i0), si, part);
                        hipLaunchKernelGGL(k_gm_fold_many, dim3((unsigned)cnt), dim3(kFold), 0, st, nbc, part, hc, si);
                        hipLaunchKernelGGL(k_gm_upd_norm<float>, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb, w, Qf,
                                           hc, cnt, si, part);
                        hipLaunchKernelGGL(k_gm_norm_fold, dim3(1), dim3(kFold), 0, st, nb, part, j, m, hc, H, cs, sn,
                                           g, si, in_tol, 0, cap);
                    } else {
                        for (int i = 0; i <= j + 1; ++i) {
                            const bool norm = i == j + 1;
                            hipLaunchKernelGGL(k_gm32_mgs, dim3((unsigned)nb), dim3(kBlock), 0, st, n, nb, w,
                                               i > 0 ? hq[i - 1] : nullptr, norm ? nullptr : hq[i], i, j, m, H, si,
                                               part);
                            if (norm)
                                hipLaunchKernelGGL(k_gm_mgs_norm, dim3(1), dim3(kFold), 0, st, nb, part, j, m, H, cs,
                                                   sn, g, si, in_tol, 0, cap);
                            else
                                hipLaunchKernelGGL(k_gm_mgs_fold, dim3(1), dim3(kFold), 0, st, nb, part, i, j, m, H,
                                                   si);
                        }
                    }
                    hipLaunchKernelGGL(k_gm32_newq, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, w, hq[j + 1], invd, z,
                                       si);
                    RSD_HIP(hipGetLastError());
                }
                hipLaunchKernelGGL(k_gm_backsub, dim3(1), dim3(1), 0, st, m, H, g, y, si);
                hipLaunchKernelGGL(k_gm32_update_d, dim3((unsigned)nbe), dim3(kBlock), 0, st, n,
                                   (const float* const*)dQ, y, invd, d, si);
                RSD_HIP(hipGetLastError());
            };
            Graph cyc;
            for (int32_t first = 1;; first = 0) {
                RSD_LAUNCH_COLS(S.mode(), k_gm32_inres, dim3((unsigned)nb), dim3(kBlock), st, n, nb, S.cols(),
                                A->val32f.p, r64, 1.0 / res, d, first, hq[0], part);
                hipLaunchKernelGGL(k_gm_residual_fold, dim3(1), dim3(kFold), 0, st, nb, part, si, first);
                if (!first) {  // a restart: the true inner residual may already meet the tolerance
                    read(si);
                    if (h->beta / (h->res0 + 1e-50) < in_tol) break;
                }
                if (first) {
                    cycle();
                } else {
                    if (!cyc.e) capture(st, cyc, cycle);
                    RSD_HIP(hipGraphLaunch(cyc.e, st));
                }
                read(si);
                if (h->stop) break;  // tolerance, breakdown or the iteration cap
            }
            total += h->total;
            ++cycles;
            hipLaunchKernelGGL(k_gm32_update_x, dim3((unsigned)nbe), dim3(kBlock), 0, st, n, x, d, res);
            res = outer_residual(0);
        }
        out.iterations = total;
        out.cycles = cycles;
        out.final_residual = res;
        out.final_residual_rel = res / (res0 + 1e-50);
    } catch (...) {
        (void)hipHostFree(h);
        throw;
    }
    (void)hipHostFree(h);
    return out;
}

}  // namespace rsd

using namespace rsd;

extern "C" int rsd_gmres_solve_mixed(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol,
                                     int64_t max_iters, int32_t krylov, int32_t precond, double inner_floor,
                                     rsd_report* rep, int64_t* outer_cycles) {
    return guard([&] {
        RSD_CHECK(ctx && A && b && x, "rsd_gmres_solve_mixed: null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(krylov >= 1 && krylov <= 512, "krylov dimension outside [1, 512]");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        const MixedGMResult r = gmres_mixed_run(ctx, A, b, x, tol, max_iters, krylov, precond,
                                                inner_floor > 0 ? inner_floor : kMixedDefaultInnerFloor);
        if (rep) {
            rep->converged = r.converged;
            rep->iterations = r.iterations;
            rep->initial_residual = r.initial_residual;
            rep->final_residual = r.final_residual;
            rep->final_residual_rel = r.final_residual_rel;
        }
        if (outer_cycles) *outer_cycles = r.cycles;
    });
}

extern "C" int rsd_gmres_solve_f64(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol,
                                   int32_t tol_type, int64_t max_iters, int32_t krylov, int32_t precond,
                                   rsd_report* rep) {
    return guard([&] {
        RSD_CHECK(ctx && A && b && x, "rsd_gmres_solve_f64: null argument");
        RSD_LIVE_MATRIX(A);
        RSD_CHECK(krylov >= 1 && krylov <= 512, "krylov dimension outside [1, 512]");
        RSD_CHECK(precond == RSD_PC_IDENTITY || precond == RSD_PC_JACOBI, "unknown preconditioner");
        RSD_CHECK(tol_type == RSD_TOL_RELATIVE || tol_type == RSD_TOL_ABSOLUTE, "unknown tolerance type");
        GMRESResult r = gmres_run(ctx, A, b, x, tol, max_iters, krylov, precond, tol_type == RSD_TOL_ABSOLUTE);
        if (rep) {
            rep->converged = r.converged;
            rep->iterations = r.iterations;
            rep->initial_residual = r.initial_residual;
            rep->final_residual = r.final_residual;
            rep->final_residual_rel = r.final_residual_rel;
        }
    });
}
