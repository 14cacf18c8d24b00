// Shared definitions of the HIP implementation of the propagator path (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fts {

// Thrown inside the library; every extern "C" entry point converts it into a
// nonzero status and a message retrievable with fts_last_error().
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define FTS_HIP_CHECK(expr)                                                                     \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            throw ::fts::Error(std::string(#expr) + ": " + hipGetErrorString(_e) + " (" +       \
                               __FILE__ + ":" + std::to_string(__LINE__) + ")");                \
    } while (0)

#define FTS_REQUIRE(cond, msg)                           \
    do {                                                 \
        if (!(cond)) throw ::fts::Error(std::string(msg)); \
    } while (0)

template <typename T>
struct vec2;
template <>
struct vec2<double> {
    using type = double2;
};
template <>
struct vec2<float> {
    using type = float2;
};

// Owning device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { allocate(bytes); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), host_(o.host_), bytes_(o.bytes_), pinned_(o.pinned_) {
        o.ptr_ = o.host_ = nullptr;
        o.bytes_ = 0;
        o.pinned_ = false;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            host_ = o.host_;
            bytes_ = o.bytes_;
            pinned_ = o.pinned_;
            o.ptr_ = o.host_ = nullptr;
            o.bytes_ = 0;
            o.pinned_ = false;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void allocate(size_t bytes) {
        release();
        if (bytes) FTS_HIP_CHECK(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    // Pinned host memory mapped into the device's address space: kernels read and
    // write it in place over the host link (the reduce-memory storage).
    void allocate_host_pinned(size_t bytes) {
        release();
        if (bytes) {
            FTS_HIP_CHECK(hipHostMalloc(&host_, bytes, hipHostMallocMapped | hipHostMallocPortable));
            FTS_HIP_CHECK(hipHostGetDevicePointer(&ptr_, host_, 0));
            pinned_ = true;
        }
        bytes_ = bytes;
    }
    void release() {
        if (pinned_) {
            if (host_) (void)hipHostFree(host_);
        } else if (ptr_) {
            (void)hipFree(ptr_);
        }
        ptr_ = host_ = nullptr;
        bytes_ = 0;
        pinned_ = false;
    }
    bool pinned_host() const { return pinned_; }
    template <typename T>
    T* as() const {
        return static_cast<T*>(ptr_);
    }
    void* get() const { return ptr_; }
    // host address of a pinned allocation (nullptr for device memory)
    void* host() const { return host_; }
    size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;    // device-usable address
    void* host_ = nullptr;   // host address of a pinned allocation
    size_t bytes_ = 0;
    bool pinned_ = false;
};

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace fts
