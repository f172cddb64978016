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

// Fixed device memory handed out in 256-byte aligned pieces and never
// freed (the scratch array of the allocation-free entry point).
struct Arena {
    char* base = nullptr;
    std::size_t size = 0, used = 0;
    void* take(std::size_t bytes) {
        const std::size_t at = (used + 255) / 256 * 256;
        if (at > size || bytes > size - at)
            throw Error(1, "evaluation needs more than the " + std::to_string(size >> 20) +
                               " MiB scratch of the allocation-free entry point");
        used = at + bytes;
        return base + at;
    }
};

// Owning device allocation, or (after use_arena) a piece of an Arena that is
// not freed.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes) { resize(bytes); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept
        : ptr_(std::exchange(o.ptr_, nullptr)), bytes_(std::exchange(o.bytes_, 0)),
          arena_(std::exchange(o.arena_, nullptr)) {}
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = std::exchange(o.ptr_, nullptr);
            bytes_ = std::exchange(o.bytes_, 0);
            arena_ = std::exchange(o.arena_, nullptr);
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void resize(std::size_t bytes) {
        if (bytes <= bytes_ && ptr_) return;
        if (arena_) {
            // a larger piece; the old one stays unused until the arena is dropped
            ptr_ = arena_->take(bytes);
            bytes_ = bytes;
            return;
        }
        release();
        if (bytes == 0) return;
        KUMO_HIP(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    void release() {
        if (ptr_ && !arena_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    // Later resizes take pieces of `arena` (the buffer must be empty).
    void use_arena(Arena* arena) {
        release();
        arena_ = arena;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    std::size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    std::size_t bytes_ = 0;
    Arena* arena_ = nullptr;
};

template <typename T> DeviceBuffer upload(const T* data, std::size_t count, hipStream_t stream) {
    DeviceBuffer b(count * sizeof(T) ? count * sizeof(T) : sizeof(T));
    if (count) KUMO_HIP(hipMemcpyAsync(b.get(), data, count * sizeof(T), hipMemcpyHostToDevice, stream));
    return b;
}

} // namespace kumo
