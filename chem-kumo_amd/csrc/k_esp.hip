// Electrostatic potential (src/electrostatic_potential.cpp:231-361) in FP64:
// V(r) = sum_A Z_A / |r - R_A| - sum_pairs sum_k c_k R^0_tuv(p, Q - r),
// with the McMurchie-Davidson Hermite integrals R from the Boys function
// (src/boys_function.cpp:47-76: table + 7-term Taylor below T = 40, closed
// form and upward recursion above). One thread per point; the primitive
// pairs are grouped by L = la + lb on the host and each group runs its own
// kernel, whose recursion size is known at compile time and whose register
// budget is its own (a high-L group does not lower the occupancy of the
// low-L ones). A group's pairs are cut into slices (grid rows) so that a few
// thousand points still fill the GPU; each slice writes its partial sum per
// point and a final kernel adds the slices in a fixed order (deterministic).
#include "kernels.hpp"

#include "common.hpp"
#include "esp_common.hpp"

#include <math.h>

namespace kumo {
namespace {
using namespace esp;

constexpr int kEspThreads = 128;

// Partial sums of slice blockIdx.y of the pairs [first_pair, last_pair) of
// one L at n points: partial[blockIdx.y * n + j].
template <int L>
__global__ void __launch_bounds__(kEspThreads)
    esp_pairs_kernel(const EspPairDev* __restrict__ pairs, int first_pair, int last_pair, int per_slice,
                     const double* __restrict__ coef, const double* __restrict__ table, PointSource ps,
                     long long first, long long n, double* __restrict__ partial) {
    const long long j = blockIdx.x * static_cast<long long>(kEspThreads) + threadIdx.x;
    if (j >= n) return;
    double x, y, z;
    point_coords(ps, first + j, x, y, z);
    const int b = first_pair + static_cast<int>(blockIdx.y) * per_slice;
    const int e = min(last_pair, b + per_slice);
    double s = 0.0;
#pragma unroll(L <= 3 ? 2 : 1) // two pairs in flight where registers allow
    for (int i = b; i < e; ++i)
        s += pair_sum<L>(pairs[i], coef, table, x, y, z);
    partial[blockIdx.y * n + j] = s;
}

// out = nuclear - (sum of the slices in order).
__global__ void __launch_bounds__(kEspThreads)
    esp_finish_kernel(const double* __restrict__ partial, int nslices, const EspAtomDev* __restrict__ atoms,
                      int natoms, PointSource ps, long long first, long long n, double* __restrict__ out) {
    const long long j = blockIdx.x * static_cast<long long>(kEspThreads) + threadIdx.x;
    if (j >= n) return;
    double x, y, z;
    point_coords(ps, first + j, x, y, z);
    double nuclear = 0.0;
    for (int a = 0; a < natoms; ++a) {
        const EspAtomDev at = atoms[a];
        const double dx = at.x - x, dy = at.y - y, dz = at.z - z;
        const double d = sqrt((dx * dx + dy * dy) + dz * dz);
        if (at.charge == 0.0 || d < 1e-8) continue;
        nuclear += at.charge / d;
    }
    double total = 0.0;
    for (int sl = 0; sl < nslices; ++sl)
        total += partial[sl * n + j];
    out[j] = nuclear - total;
}

template <int L>
void launch_group(const EspPairDev* pairs, const EspGroup& g, const double* coef, const double* table,
                  const PointSource& ps, long long first, long long n, double* partial, hipStream_t stream) {
    const int per = (g.last - g.first + g.slices - 1) / g.slices;
    const dim3 grid(static_cast<unsigned>((n + kEspThreads - 1) / kEspThreads), static_cast<unsigned>(g.slices));
    esp_pairs_kernel<L><<<grid, kEspThreads, 0, stream>>>(pairs, g.first, g.last, per, coef, table, ps, first, n,
                                                          partial);
}

__global__ void boys_kernel(const double* table, const double* T, long long n, int m, double* out) {
    const long long j = blockIdx.x * 128ll + threadIdx.x;
    if (j >= n) return;
    double v[33];
    boys(table, kBoysOrders, m, T[j], m, v);
    for (int k = 0; k <= m; ++k)
        out[j * (m + 1) + k] = v[k];
}

// boys_esp<TOP> on its own (the pair loop's variant of the Boys function):
// out[j * (TOP + 1) + m] = F_m(T[j]), m = 0..TOP.
template <int TOP>
__global__ void __launch_bounds__(128) boys_esp_kernel(const double* table, const double* T, long long n, double* out) {
    const long long j = blockIdx.x * 128ll + threadIdx.x;
    if (j >= n) return;
    double v[TOP + 1];
    boys_esp<TOP>(table + kBoysOrders * kBoysPoints + TOP * kBoysPoints * kBoysRowL, T[j], v);
#pragma unroll
    for (int k = 0; k <= TOP; ++k)
        out[j * (TOP + 1) + k] = v[k];
}

} // namespace

int esp_slices(std::vector<EspGroup>& groups, long long n) {
    // About 8 waves per SIMD over the 1024 SIMDs of the GPU; at least 16
    // pairs per slice, at most 64 slices per group.
    const long long blocks = (n + kEspThreads - 1) / kEspThreads;
    const long long want = std::max<long long>(1, (4096 + blocks - 1) / std::max<long long>(1, blocks));
    int total = 0;
    for (auto& g : groups) {
        const long long np = g.last - g.first;
        g.slices = static_cast<int32_t>(std::max<long long>(1, std::min({want, 64ll, (np + 15) / 16})));
        // no empty slice: ceil(np / slices) pairs each
        const long long per = (np + g.slices - 1) / g.slices;
        g.slices = static_cast<int32_t>(std::max<long long>(1, (np + per - 1) / per));
        total += g.slices;
    }
    return total;
}

void launch_esp(const EspPairDev* pairs, const std::vector<EspGroup>& groups, const double* coef,
                const EspAtomDev* atoms, int natoms, const double* boys_table, const PointSource& ps, long long first,
                long long n, double* partial, double* out, hipStream_t stream) {
    if (n <= 0) return;
    double* dst = partial;
    int nslices = 0;
    for (const EspGroup& g : groups) {
        if (g.last <= g.first) continue;
        if (g.L < 0 || g.L > 2 * kMaxL || g.slices < 1) invalid("invalid electrostatic potential pair group");
        switch (g.L) {
        case 0: launch_group<0>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 1: launch_group<1>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 2: launch_group<2>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 3: launch_group<3>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 4: launch_group<4>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 5: launch_group<5>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 6: launch_group<6>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 7: launch_group<7>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 8: launch_group<8>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        case 9: launch_group<9>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        default: launch_group<10>(pairs, g, coef, boys_table, ps, first, n, dst, stream); break;
        }
        KUMO_HIP(hipGetLastError());
        dst += static_cast<long long>(g.slices) * n;
        nslices += g.slices;
    }
    esp_finish_kernel<<<static_cast<unsigned>((n + kEspThreads - 1) / kEspThreads), kEspThreads, 0, stream>>>(
        partial, nslices, atoms, natoms, ps, first, n, out);
    KUMO_HIP(hipGetLastError());
}

void launch_boys_esp(const double* table, const double* T, long long n, int top, double* out, hipStream_t stream) {
    if (n <= 0) return;
    const dim3 grid(static_cast<unsigned>((n + 127) / 128)), block(128);
    switch (top) {
    case 0: boys_esp_kernel<0><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 1: boys_esp_kernel<1><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 2: boys_esp_kernel<2><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 3: boys_esp_kernel<3><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 4: boys_esp_kernel<4><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 5: boys_esp_kernel<5><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 6: boys_esp_kernel<6><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 7: boys_esp_kernel<7><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 8: boys_esp_kernel<8><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 9: boys_esp_kernel<9><<<grid, block, 0, stream>>>(table, T, n, out); break;
    case 10: boys_esp_kernel<10><<<grid, block, 0, stream>>>(table, T, n, out); break;
    default: invalid("the pair-loop Boys function has orders 0 to 10");
    }
    KUMO_HIP(hipGetLastError());
}

void launch_boys(const double* table, const double* T, long long n, int m, double* out, hipStream_t stream) {
    if (n <= 0) return;
    boys_kernel<<<static_cast<unsigned>((n + 127) / 128), 128, 0, stream>>>(table, T, n, m, out);
    KUMO_HIP(hipGetLastError());
}

} // namespace kumo
