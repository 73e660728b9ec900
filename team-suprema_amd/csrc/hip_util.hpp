// hip_util.hpp -- host-side HIP helpers shared by the single-LP engine and the
// device preprocessing: error types, the checked-call macro, an owning device
// buffer and synchronous uploads.  Internal header.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_lp.hpp"
#include "suprema_pdlp.h"

namespace supr {

struct HipFailure : std::runtime_error {
    explicit HipFailure(const std::string& m) : std::runtime_error(m) {}
};
struct BadInput : std::runtime_error {
    explicit BadInput(const std::string& m) : std::runtime_error(m) {}
};

#define SUPR_HIP(expr)                                                                       \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            throw ::supr::HipFailure(std::string("HIP error at " __FILE__ ":") +            \
                                     std::to_string(__LINE__) + " (" #expr ") -- " +       \
                                     hipGetErrorString(_e));                                 \
    } while (0)

using Clock = std::chrono::steady_clock;
inline double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Owning device allocation (move-only).
class DevBuf {
   public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    DevBuf(DevBuf&& o) noexcept : p_(o.p_) { o.p_ = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept { std::swap(p_, o.p_); return *this; }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p_) (void)hipFree(p_); }
    void alloc(size_t bytes) {
        if (p_) { (void)hipFree(p_); p_ = nullptr; }
        SUPR_HIP(hipMalloc(&p_, std::max<size_t>(bytes, 256)));
    }
    void reset() { if (p_) { (void)hipFree(p_); p_ = nullptr; } }
    template <typename T> T* as() const { return static_cast<T*>(p_); }
    explicit operator bool() const { return p_ != nullptr; }

   private:
    void* p_ = nullptr;
};

// Synchronous upload of len elements: the host array may be a temporary.
template <typename T>
DevBuf upload(const T* h, size_t len, hipStream_t s) {
    DevBuf d(len * sizeof(T));
    if (len > 0) {
        SUPR_HIP(hipMemcpyAsync(d.as<void>(), h, len * sizeof(T), hipMemcpyHostToDevice, s));
        SUPR_HIP(hipStreamSynchronize(s));
    }
    return d;
}
template <typename T>
DevBuf upload(const std::vector<T>& h, hipStream_t s) { return upload(h.data(), h.size(), s); }

template <typename T>
DevBuf device_vec(size_t len) { return DevBuf(len * sizeof(T)); }

// Maps the exception in flight to a status code and the thread's error text.
inline int map_exception() {
    try {
        throw;
    } catch (const BadInput& e) {
        set_error(e.what());
        return SUPR_EINVAL;
    } catch (const HipFailure& e) {
        set_error(e.what());
        return SUPR_EDEVICE;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return SUPR_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SUPR_EINTERNAL;
    } catch (...) {
        set_error("unknown exception");
        return SUPR_EINTERNAL;
    }
}

}  // namespace supr
