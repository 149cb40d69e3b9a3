// rsd_rowdot.hpp -- the sliced-ELL row products shared by every SpMV kernel.
#pragma once

#include "rsd_internal.hpp"

namespace rsd {

// How a kernel finds the column of the j-th entry of row r (entries in CSR
// order at slots slice_ptr[r / 64] + 64 j + r % 64):
//   COL_ABS    col[slot], an absolute int32 column per slot;
//   COL_OFF16  r + off16[slot], an int16 row-relative offset per slot;
//   COL_PAT    r + poff[pid[r] * stride + j]: the row's offset list is one
//              of a few hundred patterns shared by many rows (the stencils
//              of a structured mesh), so no index is streamed per slot.
// and, for FP32 values of a structurally symmetric matrix only:
//   COL_HALF   the values are stored once per symmetric pair, the upper one
//              (column >= row), in an ELL layout: the m-th upper value of
//              row c at m * n_pad + c.  Row r's entry j has column
//              r + poff[q * stride + j] (q = pid[r], a pattern of the half
//              layout) and its value at r + pval[q * stride + j]: for a
//              column c >= r that is the row's own m-th upper value
//              (pval = m n_pad), for c < r the value row c keeps for
//              column r (pval = m' n_pad + c - r).
enum ColMode : int { COL_ABS = 0, COL_OFF16 = 1, COL_PAT = 2, COL_HALF = 3, COL_UNI = 4 };

// uniform stencil of the FP32 half storage (COL_UNI): the union of every
// row's column offsets is u_nr runs of 3 consecutive offsets (9 for hex Q1,
// 3 for quad Q1); entry 3k+i of every row has column r + u_roff[k] + i and
// value index r + u_vidx[3k+i], rows lacking the entry hold a zero value
constexpr int kUniMaxRuns = 9;

struct Cols {
    const int64_t* slice_ptr = nullptr;
    const int32_t* rp = nullptr;     // CSR row pointer: row lengths
    const int32_t* col = nullptr;    // [slots]
    const int16_t* off16 = nullptr;  // [slots]
    const uint8_t* pid = nullptr;    // [rows]
    const int32_t* poff = nullptr;   // [n_patterns * stride]
    const int32_t* plen = nullptr;   // [n_patterns]
    const int32_t* pval = nullptr;   // [n_patterns * stride] value index - row (COL_HALF)
    int32_t stride = 0;
    int32_t u_nr = 0;                        // COL_UNI runs (3 or 9)
    int32_t u_roff[kUniMaxRuns] = {};        // first column offset of each run
    int32_t u_vidx[3 * kUniMaxRuns] = {};    // value index - row of each entry
};

// Stencil layout of an FP64 copy (COL_ST).  Every row's column offsets lie
// in one stencil of np = 3 nr positions, ascending, in nr runs of three
// consecutive offsets (c - 1, c, c + 1), c = coff[k]: the 27-point hex and
// 9-point quad Q1 stencils.  Slice s (rows 64 s .. 64 s + 63) holds np * 64
// slots from s * np * 64: positions 2p and 2p + 1 of lane l form one 16-byte
// pair at 128 p + 2 l, the last (even) position sits alone at 128 p + l.  A
// row lacking a position holds a zero there, and bit b of mask[r] tells
// which positions the row has.
constexpr int kStMaxPos = 27;

struct StCols {
    const uint32_t* mask = nullptr;  // [rows]
    int32_t np = 0;                  // 27 or 9
    int32_t coff[kStMaxPos / 3] = {};
};

// Slice-union layout of a copy the GMRES products read (COL_SU): each
// 64-row slice holds the nu positions of the union of its rows' column
// offsets (ascending, at most kSuMax), zeros where a row lacks one, in
// 16-byte pairs of positions (su_slot); mask[r] has bit k set when row r
// holds position k.  Rows of one slice then share their column offsets
// (uniform loads, coalesced x gathers) even where the rows' own patterns
// alternate, as on Q2 meshes (rows of 9, 15 and 25 entries side by side).
constexpr int kSuMax = 32;

struct SuCols {
    const int64_t* ptr = nullptr;    // [slices + 1] first slot of each slice
    const int32_t* off = nullptr;    // [slices * kSuMax] column - row of each position
    const int32_t* nu = nullptr;     // [slices] positions
    const uint32_t* mask = nullptr;  // [rows]
    int64_t n_slices = 0;
};

// slot of position k of lane l in a slice of nu positions starting at base
__host__ __device__ inline int64_t su_slot(int64_t base, int k, int nu, int l) {
    return (k == nu - 1 && (nu & 1)) ? base + 128 * (nu >> 1) + l : base + 128 * (k >> 1) + 2 * l + (k & 1);
}

// slot of position b of row r (within the copy)
__host__ __device__ inline int64_t st_slot(int64_t r, int b, int np) {
    const int64_t l = r & 63;
    const int64_t base = (r >> 6) * np * 64 + 128 * (b >> 1);
    return (b == np - 1) ? base + l : base + 2 * l + (b & 1);
}

// x entry of a run's left / right neighbour from lane l -/+ 1 (DPP wave
// shifts, 32 bits at a time); lane 0 (shr) and lane 63 (shl) keep `edge`
__device__ __forceinline__ double wave_shr1(double edge, double v) {
    const int2 e = __builtin_bit_cast(int2, edge), s = __builtin_bit_cast(int2, v);
    int2 d;
    d.x = __builtin_amdgcn_update_dpp(e.x, s.x, 0x138, 0xf, 0xf, false);
    d.y = __builtin_amdgcn_update_dpp(e.y, s.y, 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(double, d);
}
__device__ __forceinline__ double wave_shl1(double edge, double v) {
    const int2 e = __builtin_bit_cast(int2, edge), s = __builtin_bit_cast(int2, v);
    int2 d;
    d.x = __builtin_amdgcn_update_dpp(e.x, s.x, 0x130, 0xf, 0xf, false);
    d.y = __builtin_amdgcn_update_dpp(e.y, s.y, 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(double, d);
}

typedef double rsd_double2 __attribute__((ext_vector_type(2)));

// COL_ST row product over the stencil copy `val`.  The row's entries are
// added in column order, each product and sum rounded on its own; a position
// the row lacks leaves the sum unchanged (selected away by the mask, so a
// non-finite x entry outside the row's stencil cannot reach it): bitwise
// row_dot_f64's result.  Per run, one x load per lane at the centre column;
// lane l takes its left and right neighbours from lanes l - 1 and l + 1 (rows
// r - 1 and r + 1 read x[r - 1 + c] and x[r + 1 + c] as their centres), and
// the two values past the wavefront's ends are wave-uniform scalar loads.
// Must run in all 64 lanes of the wavefront, rows 64 s .. 64 s + 63 in lane
// order, also the lanes past the last row (r >= n: mask 0, clamped loads,
// result 0): their centres feed the last live row's right neighbours.
// Values are loaded in two batches (runs 0..3, then 4..NR-1).
template<int NR>
__device__ __forceinline__ double row_dot_st(int64_t r, int64_t n, const StCols& S, const double* __restrict__ val,
                                             const double* __restrict__ x) {
    constexpr int NP = 3 * NR;
    const bool live = r < n;
    const int lane = (int)(r & 63);
    const uint32_t m = live ? __builtin_nontemporal_load(S.mask + r) : 0u;
    const double* sb = val + (live ? (r >> 6) : 0) * (NP * 64);
    const int ri = (int)r, nm1 = (int)n - 1;
    const int r0 = __builtin_amdgcn_readfirstlane(ri - lane);
    double v[NP], xc[NR], xl[NR], xh[NR];
    double sum = 0.0;
    auto load_vals = [&](int b0, int b1) {
#pragma unroll
        for (int b = b0; b < b1; b += 2) {
            if (b == NP - 1) {
                v[b] = __builtin_nontemporal_load(sb + 128 * (b >> 1) + lane);
            } else {
                const rsd_double2 t = __builtin_nontemporal_load((const rsd_double2*)(sb + 128 * (b >> 1)) + lane);
                v[b] = t.x;
                v[b + 1] = t.y;
            }
        }
    };
    auto load_runs = [&](int k0, int k1) {
#pragma unroll
        for (int k = k0; k < k1; ++k) {
            const int c = S.coff[k];
            xc[k] = x[min(max(ri + c, 0), nm1)];
            xl[k] = x[min(max(r0 + c - 1, 0), nm1)];
            xh[k] = x[min(max(r0 + 64 + c, 0), nm1)];
        }
    };
    auto runs = [&](int k0, int k1) {
#pragma unroll
        for (int k = k0; k < k1; ++k) {
            const double xm = wave_shr1(xl[k], xc[k]);
            const double xp = wave_shl1(xh[k], xc[k]);
            const int b = 3 * k;
            double t = sum + v[b] * xm;
            sum = ((m >> b) & 1u) ? t : sum;
            t = sum + v[b + 1] * xc[k];
            sum = ((m >> (b + 1)) & 1u) ? t : sum;
            t = sum + v[b + 2] * xp;
            sum = ((m >> (b + 2)) & 1u) ? t : sum;
        }
    };
    // each batch issues its value loads, then its x loads, then adds (the
    // scheduling barriers keep the compiler from hoisting the x loads and
    // lane shifts of later runs ahead of the value stream)
    constexpr int K1 = NR > 4 ? 4 : NR;  // first batch: runs 0..K1-1, positions 0..3 K1 - 1 (+1 to pair up)
    constexpr int B1 = (3 * K1 + 1) & ~1;
    load_vals(0, B1 < NP ? B1 : NP);
    __builtin_amdgcn_sched_barrier(0);
    load_runs(0, K1);
    __builtin_amdgcn_sched_barrier(0);
    runs(0, K1);
    if constexpr (K1 < NR) {
        __builtin_amdgcn_sched_barrier(0);
        load_vals(B1, NP);
        __builtin_amdgcn_sched_barrier(0);
        load_runs(K1, NR);
        __builtin_amdgcn_sched_barrier(0);
        runs(K1, NR);
    }
    // the sum is formed in every lane here: callers store it for rows < n
    // only, and without this the compiler sinks the value loads into that
    // branch, behind the x loads and lane shifts
    asm volatile("" : "+v"(sum));
    return sum;
}

// Matrix values and column indices are read once per product: they are
// loaded non-temporally so the stream does not evict the x entries the
// neighbouring rows gather from L2.
template<typename T>
__device__ __forceinline__ T stream_load(const T* p) {
    return __builtin_nontemporal_load(p);
}

template<int MODE>
struct RowCols {
    int len;
    const int32_t* po;
    __device__ __forceinline__ RowCols(int64_t r, const Cols& C) {
        if constexpr (MODE == COL_PAT) {
            const int p = C.pid[r];
            len = C.plen[p];
            po = C.poff + p * C.stride;
        } else {
            len = C.rp[r + 1] - C.rp[r];
            po = nullptr;
        }
    }
    __device__ __forceinline__ int col(const Cols& C, int64_t r, int64_t slot, int j) const {
        if constexpr (MODE == COL_PAT) return (int)r + po[j];
        else if constexpr (MODE == COL_OFF16) return (int)r + (int)stream_load(C.off16 + slot);
        else return stream_load(C.col + slot);
    }
};

// The j-th column of row r is r + po[j], j < len (COL_PAT).  When every
// lane of the wavefront has the same pattern (the interior of a structured
// mesh), po and len are made wave-uniform, so the offsets come through scalar
// loads once per wavefront instead of a dependent per-lane load chain.
template<typename T, typename F>
__device__ __forceinline__ T pat_rows(int64_t r, const Cols& C, F&& body) {
    const int p = C.pid[r];
    const int p0 = __builtin_amdgcn_readfirstlane(p);
    if (__all(p == p0)) return body(C.poff + p0 * C.stride, C.plen[p0]);
    return body(C.poff + p * C.stride, C.plen[p]);
}

// Slots per batch of a row product: a batch's loads are all issued before its
// products are added in order; the slots past the last whole batch follow
// one by one.  9 divides the 27-entry Q1 hex and the 9-entry Q1 quad
// stencils (measured at 128^3: 9 beats 8 by 3 % in FP64 and 10 % in FP32;
// 14 or a predicated last batch are slower).
#ifndef RSD_ROW_UNROLL
#define RSD_ROW_UNROLL 9
#endif
constexpr int kRowUnroll = RSD_ROW_UNROLL;

// y_r = sum_j val * x[col] over row r, in CSR order, each product and sum
// rounded separately: the reference's left-to-right row sum
// (linalg/operations/backend/cpu/MatrixOps.tpp:5-23) bit for bit, whatever
// the column mode and the load batch U (which only groups the loads).
template<int MODE, int U = kRowUnroll>
__device__ __forceinline__ double row_dot_f64(int64_t r, const Cols& C, const double* __restrict__ val,
                                              const double* __restrict__ x) {
    const int64_t base = C.slice_ptr[r >> 6] + (r & 63);
    auto run = [&](auto colf, int len) {
        double sum = 0.0;
        int j = 0;
        for (; j + U <= len; j += U) {
            const int64_t s0 = base + (int64_t)j * kSlice;
            int c[U];
            double v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                c[k] = colf(s0 + k * kSlice, j + k);
                v[k] = stream_load(val + s0 + k * kSlice);
            }
            double xv[U];
#pragma unroll
            for (int k = 0; k < U; ++k) xv[k] = x[c[k]];
#pragma unroll
            for (int k = 0; k < U; ++k) sum += v[k] * xv[k];
        }
        for (; j < len; ++j) {
            const int64_t s0 = base + (int64_t)j * kSlice;
            sum += stream_load(val + s0) * x[colf(s0, j)];
        }
        return sum;
    };
    if constexpr (MODE == COL_PAT) {
        return pat_rows<double>(r, C, [&](const int32_t* po, int len) {
            return run([&](int64_t, int j) { return (int)r + po[j]; }, len);
        });
    } else {
        const RowCols<MODE> rc(r, C);
        return run([&](int64_t s, int j) { return rc.col(C, r, s, j); }, rc.len);
    }
}

// COL_HALF row product, CSR order, FP32 fused multiply-adds.  Indices are
// 32-bit (the layout is built only below 2^31 slots): one add per entry for
// the column and one for the value.  The values of the lower entries are
// read a second time by this row after the row above read them for itself:
// plain (temporal) loads keep them in L2 in between.
__device__ __forceinline__ float row_dot_half(int64_t r, const Cols& C, const float* __restrict__ val,
                                              const float* __restrict__ x) {
    constexpr int U = kRowUnroll;
    const uint32_t ru = (uint32_t)r;
    auto run = [&](const int32_t* po, const int32_t* pv, int len) {
        float sum = 0.0f;
        int j = 0;
        for (; j + U <= len; j += U) {
            uint32_t c[U];
            float v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                c[k] = ru + (uint32_t)po[j + k];
                v[k] = val[ru + (uint32_t)pv[j + k]];
            }
            float xv[U];
#pragma unroll
            for (int k = 0; k < U; ++k) xv[k] = x[c[k]];
#pragma unroll
            for (int k = 0; k < U; ++k) sum = __builtin_fmaf(v[k], xv[k], sum);
        }
        for (; j < len; ++j) sum = __builtin_fmaf(val[ru + (uint32_t)pv[j]], x[ru + (uint32_t)po[j]], sum);
        return sum;
    };
    const int p = C.pid[r];
    const int p0 = __builtin_amdgcn_readfirstlane(p);
    if (__all(p == p0)) return run(C.poff + p0 * C.stride, C.pval + p0 * C.stride, C.plen[p0]);
    return run(C.poff + p * C.stride, C.pval + p * C.stride, C.plen[p]);
}

// Uniform stencil (COL_UNI): every row reads the same NR runs of 3 entries,
// so there is no pattern id and no table load: the value indices and the
// run offsets are kernel arguments, and each run's three x values come in
// one 12-byte load.  x must be readable (and finite) from x - lo to x + n +
// hi (the layout's padding): a row lacking an entry multiplies a zero value
// by an x value outside its stencil.  The row sum takes the entries in
// column order like row_dot_half, plus fmaf(0, x, sum) for the missing
// ones, which return sum unchanged (sum starts at +0 and never becomes -0):
// bitwise row_dot_half's result.
typedef float rsd_float3u __attribute__((ext_vector_type(3), aligned(4)));

template<int NR>
__device__ __forceinline__ float row_dot_uni(int64_t r, const Cols& C, const float* __restrict__ val,
                                             const float* __restrict__ x) {
    const uint32_t ru = (uint32_t)r;
    const float* xr = x + r;
    float sum = 0.0f;
#pragma unroll
    for (int k0 = 0; k0 < NR; k0 += 3) {
        rsd_float3u xv[3];
        float v[9];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xv[k] = *(const rsd_float3u*)(xr + C.u_roff[k0 + k]);
#pragma unroll
            for (int i = 0; i < 3; ++i) v[3 * k + i] = val[ru + (uint32_t)C.u_vidx[3 * (k0 + k) + i]];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sum = __builtin_fmaf(v[3 * k], xv[k].x, sum);
            sum = __builtin_fmaf(v[3 * k + 1], xv[k].y, sum);
            sum = __builtin_fmaf(v[3 * k + 2], xv[k].z, sum);
        }
    }
    return sum;
}

// FP32 values and vectors, FP32 accumulation (fused multiply-adds)
template<int MODE>
__device__ __forceinline__ float row_dot_f32(int64_t r, const Cols& C, const float* __restrict__ val,
                                             const float* __restrict__ x) {
    if constexpr (MODE == COL_UNI) return C.u_nr == 9 ? row_dot_uni<9>(r, C, val, x) : row_dot_uni<3>(r, C, val, x);
    if constexpr (MODE == COL_HALF) return row_dot_half(r, C, val, x);
    constexpr int U = kRowUnroll;
    const int64_t base = C.slice_ptr[r >> 6] + (r & 63);
    auto run = [&](auto colf, int len) {
        float sum = 0.0f;
        int j = 0;
        for (; j + U <= len; j += U) {
            const int64_t s0 = base + (int64_t)j * kSlice;
            int c[U];
            float v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                c[k] = colf(s0 + k * kSlice, j + k);
                v[k] = stream_load(val + s0 + k * kSlice);
            }
            float xv[U];
#pragma unroll
            for (int k = 0; k < U; ++k) xv[k] = x[c[k]];
#pragma unroll
            for (int k = 0; k < U; ++k) sum = __builtin_fmaf(v[k], xv[k], sum);
        }
        for (; j < len; ++j) {
            const int64_t s0 = base + (int64_t)j * kSlice;
            sum = __builtin_fmaf(stream_load(val + s0), x[colf(s0, j)], sum);
        }
        return sum;
    };
    if constexpr (MODE == COL_PAT) {
        return pat_rows<float>(r, C, [&](const int32_t* po, int len) {
            return run([&](int64_t, int j) { return (int)r + po[j]; }, len);
        });
    } else {
        const RowCols<MODE> rc(r, C);
        return run([&](int64_t s, int j) { return rc.col(C, r, s, j); }, rc.len);
    }
}

// Row r's product over the slice-union copy, its entries added in CSR
// order (ascending positions, the mask skipping the ones the row lacks):
// FP64, each product and sum rounded separately (row_dot_f64's bits), or
// FP32 fused multiply-adds (row_dot_f32's).  Every lane of the wavefront
// must belong to the same slice, s < n_slices (the caller returns for
// wavefronts past the last slice); lanes past the last row return 0.  U
// positions (even) per batch of loads: values in pairs, then the x entries.
template<typename T>
__device__ __forceinline__ T su_madd(T sum, T v, T x) {
    if constexpr (sizeof(T) == 8) return sum + v * x;
    else return __builtin_fmaf(v, x, sum);
}

template<typename T, int U = 6>
__device__ __forceinline__ T row_dot_su(int64_t r, int64_t n, const SuCols& S, const T* __restrict__ val,
                                        const T* __restrict__ x) {
    static_assert(U % 2 == 0, "positions come in pairs");
    typedef T T2 __attribute__((ext_vector_type(2)));
    const int64_t s = __builtin_amdgcn_readfirstlane((int)(r >> 6));
    const int lane = (int)(r & 63);
    const bool live = r < n;
    const int nu = S.nu[s];
    const int32_t* off = S.off + s * kSuMax;
    const T* base = val + S.ptr[s];
    const uint32_t m = live ? S.mask[r] : 0u;
    const int ri = (int)r, nm1 = (int)n - 1;
    T sum = T(0);
    const int nu2 = nu & ~1;  // paired positions; an odd last one is single
    int k = 0;
    for (; k + U <= nu2; k += U) {
        T v[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; u += 2) {
            const T2 t = __builtin_nontemporal_load((const T2*)(base + 128 * ((k + u) >> 1)) + lane);
            v[u] = t.x;
            v[u + 1] = t.y;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[min(max(ri + off[k + u], 0), nm1)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T t = su_madd(sum, v[u], xv[u]);
            sum = ((m >> (k + u)) & 1u) ? t : sum;
        }
    }
    for (; k < nu; ++k) {
        const T v = __builtin_nontemporal_load(base + (su_slot(0, k, nu, lane)));
        const T xk = x[min(max(ri + off[k], 0), nm1)];
        const T t = su_madd(sum, v, xk);
        sum = ((m >> k) & 1u) ? t : sum;
    }
    return sum;
}

}  // namespace rsd
