// Electrostatic potential (src/electrostatic_potential.cpp:231-361) in FP64:
// V(r) = sum_A Z_A / |r - R_A| - sum_pairs sum_k c_k R^0_tuv(p, Q - r),
// with the McMurchie-Davidson Hermite integrals R from the Boys function
// (src/boys_function.cpp:47-76: table + 7-term Taylor below T = 40, closed
// form and upward recursion above). One thread per point; the primitive
// pairs are grouped by L = la + lb on the host so that each group runs a
// recursion whose size is known at compile time and stays in registers.
#include "kernels.hpp"

#include "common.hpp"

#include <math.h>

namespace kumo {
namespace {

constexpr int kBoysOrders = 39, kBoysPoints = 801;

__device__ inline void boys(const double* __restrict__ table, double T, int top, double* v) {
    const double decay = exp(-T);
    if (T >= 40.0) {
        v[0] = 0.5 * sqrt(3.141592653589793 / T);
        for (int m = 0; m < top; ++m)
            v[m + 1] = ((2.0 * m + 1.0) * v[m] - decay) / (2.0 * T);
        return;
    }
    const int point = static_cast<int>(llround(T / 0.05));
    const double delta = T - point * 0.05;
    const double* row = table + point * kBoysOrders;
    double sum = 0.0, factor = 1.0;
    for (int k = 0; k < 7; ++k) {
        sum += row[top + k] * factor;
        factor *= -delta / (k + 1.0);
    }
    v[top] = sum;
    for (int m = top; m > 0; --m)
        v[m - 1] = (2.0 * T * v[m] + decay) / (2.0 * m - 1.0);
}

__host__ __device__ constexpr int hcount(int s) { return (s + 1) * (s + 2) * (s + 3) / 6; }
// Position of (t, u, v) in the order: by t + u + v, then t descending, then u descending.
__host__ __device__ constexpr int hindex(int t, int u, int v) {
    const int s = t + u + v;
    int pos = hcount(s - 1 >= 0 ? s - 1 : -1);
    if (s == 0) return 0;
    for (int tt = s; tt > t; --tt)
        pos += s - tt + 1;
    return pos + (s - t - u);
}

template <int L>
__device__ inline double pair_sum(const EspPairDev& pr, const double* __restrict__ coef,
                                  const double* __restrict__ table, double x, double y, double z) {
    constexpr int NH = hcount(L);
    const double pc[3] = {pr.qx - x, pr.qy - y, pr.qz - z};
    double seeds[L + 1];
    boys(table, pr.p * ((pc[0] * pc[0] + pc[1] * pc[1]) + pc[2] * pc[2]), L, seeds);
    double power = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) {
        power *= -2.0 * pr.p;
        seeds[n] *= power;
    }
    double a[NH], b[NH];
    double* prev = a;
    double* cur = b;
#pragma unroll
    for (int n = L; n >= 0; --n) {
        double* tmp = prev;
        prev = cur;
        cur = tmp;
        cur[0] = seeds[n];
#pragma unroll
        for (int s = 1; s <= L - n; ++s)
#pragma unroll
            for (int t = s; t >= 0; --t)
#pragma unroll
                for (int u = s - t; u >= 0; --u) {
                    const int v = s - t - u;
                    double value;
                    if (t > 0) {
                        value = pc[0] * prev[hindex(t - 1, u, v)];
                        if (t > 1) value += (t - 1) * prev[hindex(t - 2, u, v)];
                    } else if (u > 0) {
                        value = pc[1] * prev[hindex(t, u - 1, v)];
                        if (u > 1) value += (u - 1) * prev[hindex(t, u - 2, v)];
                    } else {
                        value = pc[2] * prev[hindex(t, u, v - 1)];
                        if (v > 1) value += (v - 1) * prev[hindex(t, u, v - 2)];
                    }
                    cur[hindex(t, u, v)] = value;
                }
    }
    const double* c = coef + pr.first;
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < NH; ++k)
        sum += c[k] * cur[k];
    return sum;
}

template <int L>
__device__ double group_sum(const EspPairDev* pairs, int first, int last, const double* coef, const double* table,
                            double x, double y, double z) {
    double s = 0.0;
    for (int i = first; i < last; ++i)
        s += pair_sum<L>(pairs[i], coef, table, x, y, z);
    return s;
}

__global__ void __launch_bounds__(128) esp_kernel(const EspPairDev* pairs, int npairs, const double* coef,
                                                  const EspAtomDev* atoms, int natoms, const double* table,
                                                  PointSource ps, long long first, long long n, double* out) {
    const long long j = blockIdx.x * 128ll + threadIdx.x;
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
    int i = 0;
    while (i < npairs) {
        const int L = pairs[i].L;
        int e = i;
        while (e < npairs && pairs[e].L == L)
            ++e;
        switch (L) {
        case 0: total += group_sum<0>(pairs, i, e, coef, table, x, y, z); break;
        case 1: total += group_sum<1>(pairs, i, e, coef, table, x, y, z); break;
        case 2: total += group_sum<2>(pairs, i, e, coef, table, x, y, z); break;
        case 3: total += group_sum<3>(pairs, i, e, coef, table, x, y, z); break;
        case 4: total += group_sum<4>(pairs, i, e, coef, table, x, y, z); break;
        case 5: total += group_sum<5>(pairs, i, e, coef, table, x, y, z); break;
        case 6: total += group_sum<6>(pairs, i, e, coef, table, x, y, z); break;
        case 7: total += group_sum<7>(pairs, i, e, coef, table, x, y, z); break;
        case 8: total += group_sum<8>(pairs, i, e, coef, table, x, y, z); break;
        case 9: total += group_sum<9>(pairs, i, e, coef, table, x, y, z); break;
        default: total += group_sum<10>(pairs, i, e, coef, table, x, y, z); break;
        }
        i = e;
    }
    out[j] = nuclear - total;
}

__global__ void boys_kernel(const double* table, const double* T, long long n, int m, double* out) {
    const long long j = blockIdx.x * 128ll + threadIdx.x;
    if (j >= n) return;
    double v[33];
    boys(table, T[j], m, v);
    for (int k = 0; k <= m; ++k)
        out[j * (m + 1) + k] = v[k];
}

} // namespace

void launch_esp(const EspPairDev* pairs, int npairs, const double* coef, const EspAtomDev* atoms, int natoms,
                const double* boys_table, const PointSource& ps, long long first, long long n, double* out,
                hipStream_t stream) {
    if (n <= 0) return;
    esp_kernel<<<static_cast<unsigned>((n + 127) / 128), 128, 0, stream>>>(pairs, npairs, coef, atoms, natoms,
                                                                          boys_table, ps, first, n, out);
    KUMO_HIP(hipGetLastError());
}

void launch_boys(const double* table, const double* T, long long n, int m, double* out, hipStream_t stream) {
    if (n <= 0) return;
    boys_kernel<<<static_cast<unsigned>((n + 127) / 128), 128, 0, stream>>>(table, T, n, m, out);
    KUMO_HIP(hipGetLastError());
}

} // namespace kumo
