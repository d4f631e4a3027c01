// steppe_amd/csrc/common.hpp — shared definitions for the HIP engine.
//
// Error model (mirrors the reference's src/device/cuda/check.cuh): every HIP
// runtime call is checked; a failure throws stp::DeviceError carrying the call
// site.  The C ABI layer (capi.cpp) catches every exception at the boundary,
// records the message in a thread-local string (stp_last_error) and returns a
// status code.  Invalid arguments throw stp::ArgError.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace stp {

struct DeviceError : std::runtime_error {
    explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};
struct OomError : DeviceError {
    explicit OomError(const std::string& m) : DeviceError(m) {}
};
struct ArgError : std::invalid_argument {
    explicit ArgError(const std::string& m) : std::invalid_argument(m) {}
};

[[noreturn]] inline void hip_fail(hipError_t e, const char* what, const char* file, int line) {
    std::string msg = std::string(what) + " failed: " + hipGetErrorString(e) + " at " + file +
                      ":" + std::to_string(line);
    if (e == hipErrorOutOfMemory) throw OomError(msg);
    throw DeviceError(msg);
}

#define STP_HIP(call)                                                       \
    do {                                                                    \
        hipError_t stp_e_ = (call);                                         \
        if (stp_e_ != hipSuccess) ::stp::hip_fail(stp_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define STP_LAUNCH_CHECK() STP_HIP(hipGetLastError())

#define STP_REQUIRE(cond, msg)                                              \
    do {                                                                    \
        if (!(cond)) throw ::stp::ArgError(std::string(msg));               \
    } while (0)

inline std::int64_t ceil_div(std::int64_t a, std::int64_t b) { return (a + b - 1) / b; }
inline std::int64_t round_up(std::int64_t a, std::int64_t b) { return ceil_div(a, b) * b; }

// RAII device buffer (hipMalloc / hipFree), movable, not copyable.
template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(std::size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(std::size_t n) {
        release();
        if (n == 0) return;
        STP_HIP(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
        n_ = n;
    }
    // Grow-only reuse: keeps the allocation when it is already large enough.
    void ensure(std::size_t n) { if (n > n_) alloc(n); }
    void release() noexcept {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const noexcept { return p_; }
    std::size_t size() const noexcept { return n_; }
    std::size_t bytes() const noexcept { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

}  // namespace stp
