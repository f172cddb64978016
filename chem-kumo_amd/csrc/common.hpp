// Error handling and device memory ownership shared by the host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace kumo {

// Errors carry the C-ABI status they map to (include/kumo_hip.h).
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void invalid(const std::string& what) { throw Error(1, what); }
inline void out_of_range(const std::string& what) { throw Error(2, what); }
inline void runtime(const std::string& what) { throw Error(3, what); }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(4, std::string(what) + ": " + hipGetErrorString(e));
}
#define KUMO_HIP(call) ::kumo::hip_check((call), #call)

// Owning device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes) { resize(bytes); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) {}
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = std::exchange(o.ptr_, nullptr);
            bytes_ = std::exchange(o.bytes_, 0);
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void resize(std::size_t bytes) {
        if (bytes <= bytes_ && ptr_) return;
        release();
        if (bytes == 0) return;
        KUMO_HIP(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    std::size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    std::size_t bytes_ = 0;
};

template <typename T> DeviceBuffer upload(const T* data, std::size_t count, hipStream_t stream) {
    DeviceBuffer b(count * sizeof(T) ? count * sizeof(T) : sizeof(T));
    if (count) KUMO_HIP(hipMemcpyAsync(b.get(), data, count * sizeof(T), hipMemcpyHostToDevice, stream));
    return b;
}

} // namespace kumo
