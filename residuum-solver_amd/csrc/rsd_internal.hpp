// rsd_internal.hpp -- shared internals of libresiduum_hip.so (gfx950).
//
// Context, error plumbing, the deterministic block-tree reduction every
// reduction kernel uses, and the device-side expression interpreter.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "residuum_hip.h"

namespace rsd {

// ------------------------------------------------------------------------
// errors: every C entry point catches and maps to an int status
// ------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define RSD_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::rsd::Error(RSD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define RSD_CHECK(cond, msg)                                                               \
    do {                                                                                   \
        if (!(cond)) throw ::rsd::Error(RSD_ERR_ARG, msg);                                 \
    } while (0)

// a matrix whose problem has been destroyed (rsd_problem_destroy orphans it)
#define RSD_LIVE_MATRIX(m) RSD_CHECK((m)->pb != nullptr, "the matrix's problem has been destroyed")

template<typename F>
int guard(F&& f) {
    try {
        f();
        return RSD_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return RSD_ERR_INTERNAL;
    }
}

// ------------------------------------------------------------------------
// launch geometry
// ------------------------------------------------------------------------
constexpr int kBlock = 256;       // threads per block, 4 wave64
constexpr int kSlice = 64;        // SELL slice height = one wavefront
constexpr int kGatherMaxRow = 64;    // widest row the row gather assembly accumulates (LDS: 32 KB per wavefront)
constexpr int kGatherMaxElems = 16;  // most elements around a node the row gather assembly takes
constexpr int kXcds = 8;          // MI355X: 8 XCDs, blocks dispatched round-robin

inline int64_t nblocks(int64_t n) { return (n + kBlock - 1) / kBlock; }

// Blocks are handed to XCDs round-robin (block b -> XCD b % 8).  Remap so
// each XCD walks one contiguous range of logical blocks: neighbouring rows
// (and the x entries their stencils share) stay in one XCD's L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
    if (nb < kXcds) return b;
    const int64_t q = nb / kXcds, r = nb % kXcds;
    const int64_t x = b % kXcds, i = b / kXcds;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// ------------------------------------------------------------------------
// deterministic reduction
//   block: the terms of 256 consecutive indices (one term per thread of a
//   256-thread block) go through four 64-lane trees (lane l += lane l+off,
//   off = 32..1); thread 0 adds the four wave sums left to right and stores
//   the block sum with a plain store;
//   fold: after the kernel boundary one kFold-thread block lets lane k
//   (k < 1024) add block sums k, k+1024, ... from 0.0; the 1024 lane sums
//   go through sixteen 64-lane trees whose sums are added left to right.
// No block waits on or signals another: on the 8-XCD part a device-scope
// release/acquire pair per block writes back and invalidates the XCD's L2
// (buffer_wbl2 / buffer_inv), and having every consumer block fold the
// partials itself makes thousands of blocks read the same lines; a one-block
// fold kernel costs less than either.  The result depends only on the
// partition into 256-index blocks, never on the launch order or on which
// block finishes first.
// ------------------------------------------------------------------------
constexpr int kFold = 1024;  // threads of a fold kernel

template<typename T>
__device__ __forceinline__ T wave_tree(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off, 64);
    return v;
}

// Block sums of NV values per thread (256-thread block), stored at
// partials[k * nb + lb].
template<int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], int64_t lb, int64_t nb, double* partials) {
    __shared__ double sm[NV * 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double w = wave_tree(v[k]);
        if (lane == 0) sm[k * 4 + wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double r = sm[k * 4 + 0];
            r += sm[k * 4 + 1];
            r += sm[k * 4 + 2];
            r += sm[k * 4 + 3];
            partials[k * nb + lb] = r;
        }
    }
}

// Totals of NV arrays of block sums (array k: nbk[k] values at pk[k]); call
// from every thread of one kFold-thread block, result valid in all threads.
// Lane t adds values t, t + 1024, ... of each array from 0.0 in increasing
// order; the loads of all arrays are issued before the adds, so the fold
// costs one memory round trip per 16 Ki values.
template<int NV>
__device__ __forceinline__ void fold_partials(const double* const (&pk)[NV], const int64_t (&nbk)[NV],
                                              double (&out)[NV]) {
    __shared__ double sm[NV * 16];
    __shared__ double res[NV];
    constexpr int U = 16;
    const int t = threadIdx.x;
    double acc[NV];
    int64_t nmax = 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        acc[k] = 0.0;
        nmax = nbk[k] > nmax ? nbk[k] : nmax;
    }
    for (int64_t b0 = t; b0 < nmax; b0 += (int64_t)U * kFold) {
        double v[NV][U];
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t b = b0 + (int64_t)u * kFold;
                v[k][u] = b < nbk[k] ? pk[k][b] : 0.0;
            }
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (b0 + (int64_t)u * kFold < nbk[k]) acc[k] += v[k][u];
    }
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double w = wave_tree(acc[k]);
        if (lane == 0) sm[k * 16 + wave] = w;
    }
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double r = sm[k * 16];
#pragma unroll
            for (int i = 1; i < 16; ++i) r += sm[k * 16 + i];
            res[k] = r;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = res[k];
}

// NV arrays of nb block sums each, array k at partials + k nb
template<int NV>
__device__ __forceinline__ void fold_partials(int64_t nb, const double* __restrict__ partials, double (&out)[NV]) {
    const double* pk[NV];
    int64_t nbk[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        pk[k] = partials + k * nb;
        nbk[k] = nb;
    }
    fold_partials<NV>(pk, nbk, out);
}

// ------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------
struct Scratch {
    double* partials = nullptr;   // reduction partials
    int64_t partials_cap = 0;     // doubles
};

}  // namespace rsd

struct rsd_ctx {
    int device = 0;
    int n_cu = 256;               // compute units of the device
    hipStream_t stream = nullptr;
    rsd::Scratch scratch;
    double* d_scalars = nullptr;  // small device scalar area for BLAS-1 results
    int32_t gmres_orth = RSD_ORTHO_MGS;  // GMRES orthogonalisation (rsd_ctx_set_gmres_orthogonalization)
    // make sure the partials area can hold `need` doubles
    double* partials(int64_t need);
    // named device workspaces, grown on demand and kept for the context's life
    struct Ws { void* p = nullptr; size_t bytes = 0; };
    std::vector<std::pair<std::string, Ws>> ws_tab;
    void* ws(const std::string& name, size_t bytes);
    ~rsd_ctx();
};

namespace rsd {

// ------------------------------------------------------------------------
// expressions (compiled on the host, interpreted on the device)
// ------------------------------------------------------------------------
constexpr int kExprMax = 48;
enum ExprOp : int32_t { E_CONST, E_VAR, E_ADD, E_SUB, E_MUL, E_DIV, E_POW, E_NEG,
                        E_SIN, E_COS, E_TAN, E_EXP, E_LOG, E_SQRT, E_ABS, E_MIN, E_MAX };
struct Expr {
    int32_t n = 0;
    int32_t uses_t = 0;
    int32_t uses_T = 0;
    int32_t op[kExprMax];
    double c[kExprMax];
};
Expr compile_expr(const std::string& text);
double eval_expr_host(const Expr& e, double t, double x, double y, double z, double T);

__device__ inline double eval_expr(const Expr& e, double t, double x, double y, double z, double T) {
    double st[16];
    int sp = 0;
    for (int i = 0; i < e.n; ++i) {
        switch (e.op[i]) {
        case E_CONST: st[sp++] = e.c[i]; break;
        case E_VAR: {
            const int v = (int)e.c[i];
            st[sp++] = v == 0 ? t : v == 1 ? x : v == 2 ? y : v == 3 ? z : T;
            break;
        }
        case E_ADD: sp--; st[sp - 1] = st[sp - 1] + st[sp]; break;
        case E_SUB: sp--; st[sp - 1] = st[sp - 1] - st[sp]; break;
        case E_MUL: sp--; st[sp - 1] = st[sp - 1] * st[sp]; break;
        case E_DIV: sp--; st[sp - 1] = st[sp - 1] / st[sp]; break;
        case E_POW: sp--; st[sp - 1] = pow(st[sp - 1], st[sp]); break;
        case E_MIN: sp--; st[sp - 1] = (st[sp] < st[sp - 1]) ? st[sp] : st[sp - 1]; break;
        case E_MAX: sp--; st[sp - 1] = (st[sp] > st[sp - 1]) ? st[sp] : st[sp - 1]; break;
        case E_NEG: st[sp - 1] = -st[sp - 1]; break;
        case E_SIN: st[sp - 1] = sin(st[sp - 1]); break;
        case E_COS: st[sp - 1] = cos(st[sp - 1]); break;
        case E_TAN: st[sp - 1] = tan(st[sp - 1]); break;
        case E_EXP: st[sp - 1] = exp(st[sp - 1]); break;
        case E_LOG: st[sp - 1] = log(st[sp - 1]); break;
        case E_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
        default: st[sp - 1] = fabs(st[sp - 1]); break;
        }
    }
    return st[0];
}

// an instantiated hipGraph of work recorded on a stream (kernel launches
// only: a memset node is never captured)
struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    Graph() = default;
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    ~Graph() { reset(); }
    void reset() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
        e = nullptr;
        g = nullptr;
    }
};

// records body()'s launches on st into G (replacing what G held)
template<typename F>
void capture(hipStream_t st, Graph& G, F&& body) {
    G.reset();
    RSD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(st, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    RSD_HIP(hipStreamEndCapture(st, &G.g));
    RSD_HIP(hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0));
}

// ------------------------------------------------------------------------
// device buffer helper
// ------------------------------------------------------------------------
template<typename T>
struct DBuf {
    T* p = nullptr;
    int64_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(int64_t count) {
        reset();
        if (count > 0) RSD_HIP(hipMalloc(&p, sizeof(T) * (size_t)count));
        n = count;
    }
    void upload(const T* h, int64_t count, hipStream_t s) {
        alloc(count);
        if (count > 0) RSD_HIP(hipMemcpyAsync(p, h, sizeof(T) * (size_t)count, hipMemcpyHostToDevice, s));
    }
};

}  // namespace rsd
