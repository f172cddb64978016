// Device functions of the exact ESP pair sums shared by the exact and the
// far-field ESP kernels: the pair loop's Boys function, the order of the
// Hermite indices and the sum over one pair's Hermite terms at a point.
#pragma once

#include "kernels.hpp"

#include <math.h>

namespace kumo {
namespace esp {

// F_0..F_top(T) into v[0..top] in the reference's arithmetic
// (src/boys_function.cpp:47-76). Row i of `table` (stride doubles apart)
// holds F_top .. F_top+6 at i * 0.05 from offset `first`.
__device__ __attribute__((always_inline)) inline void boys(const double* __restrict__ table, int stride, int first,
                                                           double T, int top, double* v) {
    const double decay = exp(-T);
    if (T >= 40.0) {
        v[0] = 0.5 * sqrt(3.141592653589793 / T);
#pragma unroll
        for (int m = 0; m < top; ++m)
            v[m + 1] = ((2.0 * m + 1.0) * v[m] - decay) / (2.0 * T);
        return;
    }
    const int point = static_cast<int>(llround(T / 0.05));
    const double delta = T - point * 0.05;
    const double* row = table + point * stride + first;
    double sum = 0.0, factor = 1.0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        sum += row[k] * factor;
        factor *= -delta / (k + 1.0);
    }
    v[top] = sum;
#pragma unroll
    for (int m = top; m > 0; --m)
        v[m - 1] = (2.0 * T * v[m] + decay) / (2.0 * m - 1.0);
}

// boys() for the pair loop of the ESP kernels: the same scheme (the table
// point nearest to T and a 7-term Taylor series below T = 40, the closed form
// and upward recursion above), with every division by a constant, and the
// repeated division by 2T, replaced by a product with the reciprocal and the
// series summed by Horner's rule, and above T = 40 the reciprocal square
// root refined by Newton steps. The values agree with boys() to a few ulp;
// the loop then needs no FP64 division or square root (boys(): up to 7 + L
// divisions and a square root).
template <int TOP>
__device__ __attribute__((always_inline)) inline void boys_esp(const double* __restrict__ rows, double T,
                                                               double (&v)[TOP + 1]) {
    constexpr double kInvFactorial[7] = {1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0};
    if (T >= 40.0) {
        // 1/sqrt(T) from the hardware estimate and two Newton steps (each
        // squares the relative error), then F_0 = (sqrt(pi) / 2) / sqrt(T)
        // and 1 / (2T) without a division or a square root.
        double y = __builtin_amdgcn_rsq(T);
        y = y * (1.5 - 0.5 * T * y * y);
        y = y * (1.5 - 0.5 * T * y * y);
        v[0] = 0.88622692545275801 * y;
        if constexpr (TOP > 0) {
            const double decay = exp(-T), half = 0.5 * (y * y);
#pragma unroll
            for (int m = 0; m < TOP; ++m)
                v[m + 1] = ((2.0 * m + 1.0) * v[m] - decay) * half;
        }
        return;
    }
    const int point = static_cast<int>(llround(T * 20.0)); // nearest table point (step 0.05)
    const double x = point * 0.05 - T;
    const double* row = rows + point * kBoysRowL;
    double sum = row[6] * kInvFactorial[6];
#pragma unroll
    for (int k = 5; k >= 0; --k)
        sum = sum * x + row[k] * kInvFactorial[k];
    v[TOP] = sum;
    if constexpr (TOP > 0) {
        const double decay = exp(-T), t2 = 2.0 * T;
#pragma unroll
        for (int m = TOP; m > 0; --m)
            v[m - 1] = (t2 * v[m] + decay) * (1.0 / (2.0 * m - 1.0));
    }
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

// sum_k c_k R^0_k of one pair. The recursion from order n + 1 to order n of
// the auxiliary index runs in place over one array: the entries of total
// degree s at order n need those of degrees s - 1 and s - 2 at order n + 1,
// so the degrees are rewritten from the highest down.
template <int L>
__device__ __attribute__((always_inline)) inline double pair_sum(const EspPairDev& pr, const double* __restrict__ coef,
                                                                 const double* __restrict__ table, double x, double y,
                                                                 double z) {
    constexpr int NH = hcount(L);
    const double pc[3] = {pr.qx - x, pr.qy - y, pr.qz - z};
    double seeds[L + 1];
    // the orders L .. L + 6 of this L (kBoysRowL doubles per point)
    boys_esp<L>(table + kBoysOrders * kBoysPoints + L * kBoysPoints * kBoysRowL,
                pr.p * ((pc[0] * pc[0] + pc[1] * pc[1]) + pc[2] * pc[2]), seeds);
    double power = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) {
        power *= -2.0 * pr.p;
        seeds[n] *= power;
    }
    double r[NH];
#pragma unroll
    for (int n = L; n >= 0; --n) {
#pragma unroll
        for (int s = L - n; s >= 1; --s)
#pragma unroll
            for (int t = s; t >= 0; --t)
#pragma unroll
                for (int u = s - t; u >= 0; --u) {
                    const int v = s - t - u;
                    double value;
                    if (t > 0) {
                        value = pc[0] * r[hindex(t - 1, u, v)];
                        if (t > 1) value += (t - 1) * r[hindex(t - 2, u, v)];
                    } else if (u > 0) {
                        value = pc[1] * r[hindex(t, u - 1, v)];
                        if (u > 1) value += (u - 1) * r[hindex(t, u - 2, v)];
                    } else {
                        value = pc[2] * r[hindex(t, u, v - 1)];
                        if (v > 1) value += (v - 1) * r[hindex(t, u, v - 2)];
                    }
                    r[hindex(t, u, v)] = value;
                }
        r[0] = seeds[n];
    }
    const double* c = coef + pr.first;
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < NH; ++k)
        sum += c[k] * r[k];
    return sum;
}

} // namespace esp
} // namespace kumo
