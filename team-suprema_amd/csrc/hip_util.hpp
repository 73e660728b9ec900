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

// Device memory behind DevBuf (pdlp_engine.hip).  Blocks a call releases
// are kept per device, up to SUPR_DEVICE_CACHE_MB (default 4096; 0 turns the
// cache off), and handed to later allocations they fit, so repeated calls
// skip hipMalloc / hipFree; supr_release_cached_memory() frees them.  A block
// is cached only after the device has finished its work (a device-wide
// synchronisation, as hipFree does), so a reused block is never still in use.
void* device_alloc(size_t bytes, size_t* reserved, int* device);
void device_free(void* p, size_t reserved, int device);

// Device allocation (move-only): owning, or a borrowed view into memory that
// something else owns (an Arena).
class DevBuf {
   public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    static DevBuf borrow(void* p) {
        DevBuf b;
        b.p_ = p;
        b.own_ = false;
        return b;
    }
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), bytes_(o.bytes_), dev_(o.dev_), own_(o.own_) { o.p_ = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(bytes_, o.bytes_);
        std::swap(dev_, o.dev_);
        std::swap(own_, o.own_);
        return *this;
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t bytes) {
        release();
        own_ = true;
        p_ = device_alloc(std::max<size_t>(bytes, 256), &bytes_, &dev_);
    }
    void reset() { release(); }
    template <typename T> T* as() const { return static_cast<T*>(p_); }
    explicit operator bool() const { return p_ != nullptr; }

   private:
    void release() {
        if (p_ && own_) device_free(p_, bytes_, dev_);
        p_ = nullptr;
    }
    void* p_ = nullptr;
    size_t bytes_ = 0;
    int dev_ = 0;
    bool own_ = true;
};

// Many device buffers carved out of one allocation (one hipMalloc and one
// hipFree for all of them instead of one each).  Used in two passes: the
// same sequence of take() calls first counts the bytes (before commit(),
// the views are null), then, after commit(), hands out the real views.
// Every buffer starts on a 256-byte boundary.
class Arena {
   public:
    template <typename T> DevBuf take(size_t len) {
        const size_t off = (used_ + 255) / 256 * 256;
        used_ = off + std::max<size_t>(len * sizeof(T), 1);
        return DevBuf::borrow(base_ ? static_cast<char*>(base_) + off : nullptr);
    }
    size_t bytes() const { return used_; }
    void commit() {
        mem_.alloc(used_);
        base_ = mem_.as<void>();
        used_ = 0;
    }

   private:
    DevBuf mem_;
    void* base_ = nullptr;
    size_t used_ = 0;
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
