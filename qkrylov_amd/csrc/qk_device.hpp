// HIP runtime helpers shared by the kernels and the host drivers.
#pragma once

#include <hip/hip_runtime.h>

#include "qk_common.hpp"

#include <string>
#include <utility>

namespace qk {

#define QK_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t qk_err_ = (expr);                                              \
        if (qk_err_ != hipSuccess)                                                \
            ::qk::fail(std::string("HIP error ") + hipGetErrorString(qk_err_) +   \
                       " at " + __FILE__ + ":" + std::to_string(__LINE__) +       \
                       " in " #expr);                                             \
    } while (0)

// Owning device allocation (hipMalloc / hipFree).
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { resize(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void resize(size_t bytes) {
        if (bytes <= n_) return;
        release();
        if (bytes) QK_HIP(hipMalloc(&p_, bytes));
        n_ = bytes;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    template <typename T = void>
    T* get() const { return static_cast<T*>(p_); }
    size_t bytes() const { return n_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

template <typename T>
inline void upload(DevBuf& d, const T* h, size_t n, hipStream_t st) {
    d.resize(sizeof(T) * (n ? n : 1));
    if (n) QK_HIP(hipMemcpyAsync(d.get(), h, sizeof(T) * n, hipMemcpyHostToDevice, st));
}

int device_cu_count();

}  // namespace qk
