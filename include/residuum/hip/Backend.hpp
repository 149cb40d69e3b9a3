// residuum/hip/Backend.hpp -- the Hip backend tag.
//
// The reference's containers are templated on a backend tag with the duck
// type alloc / copy / zero / set / Ptr (linalg/types/backend/CPU.hpp:15-40;
// the unused CUDA scaffold in linalg/types/backend/CUDA.hpp:16-73).
// backend::Hip satisfies it with device memory from libresiduum_hip.so, so
// the reference's own Vector<T, Hip> and CSRMatrix<T, Hip>
// (linalg/types/Vector.hpp:10-45, CSRMatrix.hpp:14-88) hold their data in HBM
// unchanged.  Include this header with the reference's include directory on
// the include path; link libresiduum_hip.so.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>

#include "core/Types.hpp"
#include "linalg/types/backend/cuda/CopyKind.hpp"
#include "residuum_hip.h"

namespace residuum {
namespace hip {

// C-ABI status -> exception; the reference reports backend failures with
// std::runtime_error (CUDA.hpp:32-35) and a missing diagonal with
// std::out_of_range (CSRMatrix.hpp:70)
inline void check(int rc) {
    if (rc == RSD_OK) return;
    throw std::runtime_error(std::string("residuum_hip: ") + rsd_last_error_string());
}

// The HIP context (one stream plus scratch) used by every operation on the
// calling host thread: all work is stream-ordered on it, which keeps the
// reference's synchronous call semantics for the host code around it.
class Context {
public:
    static rsd_ctx* get() { return instance().ctx_; }
    // Select the device for this thread's context; only before first use.
    static void select_device(int device) {
        if (device_slot() != device && created()) throw std::runtime_error("residuum_hip: context already created");
        device_slot() = device;
    }
    static void synchronize() { check(rsd_ctx_sync(get())); }

    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ~Context() {
        if (ctx_) rsd_ctx_destroy(ctx_);
    }

private:
    Context() { check(rsd_ctx_create(device_slot(), &ctx_)); created() = true; }
    static Context& instance() {
        thread_local Context c;
        return c;
    }
    static int& device_slot() {
        thread_local int d = 0;
        return d;
    }
    static bool& created() {
        thread_local bool c = false;
        return c;
    }
    rsd_ctx* ctx_ = nullptr;
};

namespace detail {

// CSR patterns made by Assembler<Hip>::createMatrix, by row pointer: the
// solvers then run on the device problem's own sliced layout of that
// pattern instead of rebuilding it from the CSR arrays.  An entry goes when
// either array is freed.
struct PatternOwner {
    const void* colIdx;
    rsd_problem* pb;
};
inline std::unordered_map<const void*, PatternOwner>& patterns() {
    thread_local std::unordered_map<const void*, PatternOwner> m;
    return m;
}
inline void remember_pattern(const void* rowPtr, const void* colIdx, rsd_problem* pb) {
    patterns()[rowPtr] = PatternOwner{colIdx, pb};
}
inline rsd_problem* pattern_problem(const void* rowPtr, const void* colIdx) {
    auto& m = patterns();
    auto it = m.find(rowPtr);
    return (it != m.end() && it->second.colIdx == colIdx) ? it->second.pb : nullptr;
}
inline void forget_pointer(const void* p) {
    auto& m = patterns();
    if (m.empty()) return;
    m.erase(p);
    for (auto it = m.begin(); it != m.end();) {
        if (it->second.colIdx == p) it = m.erase(it);
        else ++it;
    }
}
inline void forget_problem(const rsd_problem* pb) {
    auto& m = patterns();
    for (auto it = m.begin(); it != m.end();) {
        if (it->second.pb == pb) it = m.erase(it);
        else ++it;
    }
}

}  // namespace detail

}  // namespace hip

namespace linalg {
namespace types {
namespace backend {

template<typename T>
struct HipDeleter {
    void operator()(T* p) const noexcept {
        if (!p) return;
        hip::detail::forget_pointer(p);
        rsd_free(hip::Context::get(), p);
    }
};

class Hip {
public:
    template<typename T>
    using Ptr = std::unique_ptr<T[], HipDeleter<T>>;

    template<typename T>
    static Ptr<T> alloc(Index n) {
        void* p = nullptr;
        if (n) hip::check(rsd_malloc(hip::Context::get(), sizeof(T) * n, &p));
        return Ptr<T>(static_cast<T*>(p));
    }

    template<typename T>
    static void copy(T* dst, const T* src, Index n, CopyKind kind) {
        if (!n) return;
        int k = RSD_COPY_D2D;
        switch (kind) {
        case CopyKind::HostToHost: k = RSD_COPY_H2H; break;
        case CopyKind::HostToDevice: k = RSD_COPY_H2D; break;
        case CopyKind::DeviceToHost: k = RSD_COPY_D2H; break;
        case CopyKind::DeviceToDevice: k = RSD_COPY_D2D; break;
        }
        hip::check(rsd_memcpy(hip::Context::get(), dst, src, sizeof(T) * n, k));
    }

    template<typename T>
    static void zero(T* ptr, Index n) {
        if (n) hip::check(rsd_memset_zero(hip::Context::get(), ptr, sizeof(T) * n));
    }

    template<typename T>
    static void set(T* ptr, Index n, T value) {
        if constexpr (std::is_same_v<T, double>) hip::check(rsd_fill_f64(hip::Context::get(), ptr, (int64_t)n, value));
        else if constexpr (std::is_same_v<T, float>) hip::check(rsd_fill_f32(hip::Context::get(), ptr, (int64_t)n, value));
        else static_assert(sizeof(T) == 0, "Hip::set: double and float only");
    }
};

}  // namespace backend
}  // namespace types
}  // namespace linalg
}  // namespace residuum
