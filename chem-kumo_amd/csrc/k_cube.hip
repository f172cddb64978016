// Cube text on the GPU: the samples of Field::write_cube (src/cube.cpp:104-109)
// formatted where they lie in HBM, so that only finished text crosses PCIe.
//
// Layout of the sample section: value i takes 13 bytes (printf("%13.5e"),
// exactly 13 for every finite double), followed by a line break when
// (i + 1) % 6 == 0 or (i + 1) % nz == 0, nz the count of the third grid axis.
// The text offset of value i is therefore closed-form:
//   T(i) = 13 i + floor(i / 6) + floor(i / nz) - floor(i / lcm(6, nz)),
// and each workgroup formats 256 consecutive values into LDS and writes its
// contiguous byte range with aligned 32-bit stores (bytes at its two ends).
#include "common.hpp"
#include "kernels.hpp"
#include "sci_format.hpp"

#include <algorithm>

namespace kumo {
namespace {

__constant__ sci::Pow10 kPow10[] = {KUMO_POW10_TABLE};

constexpr int kThreads = 256;

__device__ __forceinline__ long long text_offset(long long i, long long nz, long long period) {
    return 13 * i + i / 6 + i / nz - i / period;
}

__global__ __launch_bounds__(kThreads) void cube_text_kernel(const double* __restrict__ values, long long first,
                                                             long long count, long long nz, long long period,
                                                             long long text_first, unsigned char* __restrict__ text) {
    __shared__ unsigned char buf[kThreads * 14];
    const int t = threadIdx.x;
    const long long g0 = first + static_cast<long long>(blockIdx.x) * kThreads;
    const int nv = static_cast<int>(min(static_cast<long long>(kThreads), first + count - g0));
    // Residues of the block's first value: the offsets within the block need
    // only 32-bit arithmetic (nz < 2^31; the period quotient is handled below).
    const unsigned r6 = static_cast<unsigned>(g0 % 6);
    const unsigned rz = static_cast<unsigned>(g0 % nz);
    const long long rp = g0 % period;
    const long long base = text_offset(g0, nz, period);
    const unsigned znz = static_cast<unsigned>(nz);
    if (t < nv) {
        const unsigned u = static_cast<unsigned>(t);
        const unsigned wrap_p = period < (1LL << 31) ? (static_cast<unsigned>(rp) + u) / static_cast<unsigned>(period)
                                                     : (rp + u >= period ? 1u : 0u);
        const int off = static_cast<int>(13 * u + (r6 + u) / 6 + (rz + u) / znz - wrap_p);
        char s[13];
        sci::format_cube_sample(values[g0 + t], kPow10, false, s);
        for (int c = 0; c < 13; ++c) buf[off + c] = static_cast<unsigned char>(s[c]);
        if ((r6 + u + 1) % 6 == 0 || (rz + u + 1) % znz == 0) buf[off + 13] = '\n';
    }
    __syncthreads();
    const int nbytes = static_cast<int>(text_offset(g0 + nv, nz, period) - base);
    unsigned char* dst = text + (base - text_first);
    const int head = min(nbytes, static_cast<int>((4 - (reinterpret_cast<uintptr_t>(dst) & 3)) & 3));
    const int nwords = (nbytes - head) >> 2;
    const int tail = head + 4 * nwords;
    if (t < head) dst[t] = buf[t];
    for (int w = t; w < nwords; w += kThreads) {
        const int o = head + 4 * w;
        const uint32_t word = static_cast<uint32_t>(buf[o]) | static_cast<uint32_t>(buf[o + 1]) << 8 |
                              static_cast<uint32_t>(buf[o + 2]) << 16 | static_cast<uint32_t>(buf[o + 3]) << 24;
        *reinterpret_cast<uint32_t*>(dst + o) = word;
    }
    if (t < nbytes - tail) dst[tail + t] = buf[tail + t];
}

// Fixed-width slots: printf("%*.*e", P + 8, P, v) right-justified in P + 8
// bytes per value.
__global__ __launch_bounds__(kThreads) void sci_slots_kernel(const double* __restrict__ values, long long n, int P,
                                                             int exact_only, char* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    char s[24];
    const int len = sci::format_sci(values[i], P, kPow10, exact_only != 0, s);
    const int width = P + 8;
    char* o = out + i * width;
    for (int c = 0; c < width - len; ++c) o[c] = ' ';
    for (int c = 0; c < len; ++c) o[width - len + c] = s[c];
}

__global__ __launch_bounds__(kThreads) void count_nonfinite_kernel(const double* __restrict__ values, long long n,
                                                                   int* __restrict__ bad) {
    long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kThreads;
    int local = 0;
    for (; i < n; i += stride) {
        const uint64_t b = __builtin_bit_cast(uint64_t, values[i]);
        local += ((b >> 52) & 0x7ff) == 0x7ff;
    }
    if (local) atomicAdd(bad, local);
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

} // namespace

long long cube_text_offset(long long i, long long nz) {
    const long long period = nz % 3 == 0 ? (nz % 2 == 0 ? nz : 2 * nz) : (nz % 2 == 0 ? 3 * nz : 6 * nz);
    return 13 * i + i / 6 + i / nz - i / period;
}

void launch_cube_text(const double* values, long long first, long long count, long long nz, long long text_first,
                      unsigned char* text, hipStream_t stream) {
    if (count <= 0) return;
    const long long period = nz % 3 == 0 ? (nz % 2 == 0 ? nz : 2 * nz) : (nz % 2 == 0 ? 3 * nz : 6 * nz);
    const long long blocks = (count + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(cube_text_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, stream, values, first,
                       count, nz, period, text_first, text);
    KUMO_HIP(hipGetLastError());
}

void launch_sci_slots(const double* values, long long n, int precision, int exact_only, char* out,
                      hipStream_t stream) {
    if (n <= 0) return;
    const long long blocks = (n + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(sci_slots_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, stream, values, n,
                       precision, exact_only, out);
    KUMO_HIP(hipGetLastError());
}

void launch_count_nonfinite(const double* values, long long n, int* bad, hipStream_t stream) {
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, stream, bad);
    KUMO_HIP(hipGetLastError());
    if (n <= 0) return;
    const long long blocks = std::min<long long>((n + kThreads - 1) / kThreads, 4096);
    hipLaunchKernelGGL(count_nonfinite_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, stream, values,
                       n, bad);
    KUMO_HIP(hipGetLastError());
}

} // namespace kumo
