// residuum/hip/Backend.hpp -- the Hip backend tag and device containers.
//
// C++20 counterpart of the reference's Backend-tag seam
// (linalg/types/backend/CPU.hpp:15-40, the CUDA scaffold in
// linalg/types/backend/CUDA.hpp:16-73) and of the containers that consume
// it (linalg/types/Vector.hpp:10-45, linalg/types/CSRMatrix.hpp:14-88).
// Everything calls libresiduum_hip.so through include/residuum_hip.h.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "residuum_hip.h"

namespace residuum::hip {

using Real = double;
using Index = std::size_t;

// status -> exception, as the reference throws std::runtime_error (CUDA.hpp:32-35)
inline void check(int rc) {
    if (rc != RSD_OK) throw std::runtime_error(std::string("residuum_hip: ") + rsd_last_error_string());
}

// One rsd_ctx (one HIP stream) per host thread, created on first use.
class Context {
public:
    static rsd_ctx* get(int device = 0) {
        thread_local Context c(device);
        return c.ctx_;
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ~Context() {
        if (ctx_) rsd_ctx_destroy(ctx_);
    }

private:
    explicit Context(int device) { check(rsd_ctx_create(device, &ctx_)); }
    rsd_ctx* ctx_ = nullptr;
};

// mirrors linalg/types/backend/cuda/CopyKind.hpp:9-14
enum class CopyKind { HostToHost = RSD_COPY_H2H, HostToDevice = RSD_COPY_H2D, DeviceToHost = RSD_COPY_D2H,
                      DeviceToDevice = RSD_COPY_D2D };

template<typename T>
struct HipDeleter {
    void operator()(T* p) const noexcept {
        if (p) rsd_free(Context::get(), p);
    }
};

// The backend tag: alloc / copy / zero / set, same duck type as CPU.hpp:15-40.
class Hip {
public:
    template<typename T>
    using Ptr = std::unique_ptr<T, HipDeleter<T>>;

    template<typename T>
    static Ptr<T> alloc(Index n) {
        void* p = nullptr;
        check(rsd_malloc(Context::get(), n * sizeof(T), &p));
        return Ptr<T>(static_cast<T*>(p));
    }
    template<typename T>
    static void copy(T* dst, const T* src, Index n, CopyKind kind) {
        check(rsd_memcpy(Context::get(), dst, src, n * sizeof(T), static_cast<int>(kind)));
    }
    template<typename T>
    static void zero(T* p, Index n) {
        check(rsd_memset_zero(Context::get(), p, n * sizeof(T)));
    }
    template<typename T>
    static void set(T* p, Index n, T v) {
        if constexpr (sizeof(T) == 8) check(rsd_fill_f64(Context::get(), reinterpret_cast<double*>(p), (int64_t)n, (double)v));
        else check(rsd_fill_f32(Context::get(), reinterpret_cast<float*>(p), (int64_t)n, (float)v));
    }
};

// Vector<Real, Hip> (linalg/types/Vector.hpp:10-45): move-only, device data.
template<typename T>
class Vector {
public:
    using value_type = T;
    using backend_type = Hip;

    explicit Vector(Index size) : size_(size), data_(Hip::alloc<T>(size)) {}
    Vector(const Vector&) = delete;
    Vector& operator=(const Vector&) = delete;
    Vector(Vector&&) = default;
    Vector& operator=(Vector&&) = default;

    Index size() const { return size_; }
    T* data() { return data_.get(); }
    const T* data() const { return data_.get(); }
    void zero() { Hip::zero(data_.get(), size_); }
    void set(T v) { Hip::set(data_.get(), size_, v); }

    // host transfers (the reference reads data() on the host; here it is device memory)
    void upload(const std::vector<T>& h) {
        if (h.size() != size_) throw std::runtime_error("Vector::upload: size mismatch");
        Hip::copy(data_.get(), h.data(), size_, CopyKind::HostToDevice);
    }
    std::vector<T> download() const {
        std::vector<T> h(size_);
        Hip::copy(h.data(), data_.get(), size_, CopyKind::DeviceToHost);
        return h;
    }

private:
    Index size_;
    typename Hip::Ptr<T> data_;
};

}  // namespace residuum::hip
