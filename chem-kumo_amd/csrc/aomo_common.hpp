// Helpers shared by the AO->MO contraction kernels: MFMA operand vector
// types, the moments a contraction mode makes nonzero and the splitting of
// Phi values into bf16 terms.
#pragma once

#include "kernels.hpp"

#include <hip/hip_runtime.h>

namespace kumo {
namespace aomo {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

// Moments a contraction can make nonzero: the value; + gradient and tau terms
// (order 1); + the trace (Laplacian panel), the diagonal or all six Hessian
// moments (order 2). The others are written as zeros without being summed.
template <int ORDER, unsigned PSET>
__host__ __device__ __attribute__((always_inline)) constexpr bool moment_used(int k) {
    if (k == 0) return true;
    if (ORDER == 0) return false;
    if (k <= 6) return true;
    if (ORDER == 1) return false;
    if ((PSET >> kPanelLap) & 1u) return k == 7;
    if (PSET == kPanelsDiag) return k == 7 || k == 10 || k == 12;
    return true;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Splits two FP64 values into S bf16 terms, packed in pairs (element 0 in the
// low half): t0 = bf16(v); the residual v - t0, exact in FP64, is rounded
// once to FP32 (2^-33 relative to v) and split into the remaining terms
// (each FP32 residual exact), so three terms represent v to about 2^-26. The empty asm keeps
// each FP64->FP32 rounding apart from the bf16 rounding after it (the
// compiler would fuse them into a slower direct FP64->bf16 rounding;
// rounding twice is exact here since FP32 has more than twice bf16's
// precision).
template <int S>
__device__ __attribute__((always_inline)) inline void split_pair(double v0, double v1, uint32_t (&t)[S]) {
    f32x2 f = {static_cast<float>(v0), static_cast<float>(v1)};
    asm("" : "+v"(f));
    t[0] = __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2));
    f32x2 r = {static_cast<float>(v0 - static_cast<double>(__uint_as_float(t[0] << 16))),
               static_cast<float>(v1 - static_cast<double>(__uint_as_float(t[0] & 0xffff0000u)))};
    asm("" : "+v"(r));
#pragma unroll
    for (int k = 1; k < S; ++k) {
        t[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
        if (k + 1 < S) r = f32x2{r.x - __uint_as_float(t[k] << 16), r.y - __uint_as_float(t[k] & 0xffff0000u)};
    }
}

// Splits two FP32 values into S bf16 terms, packed in pairs (element 0 in
// the low half): each term is the bf16 rounding of what the previous terms
// leave, every residual exact in FP32; three terms carry the 24 bits.
template <int S>
__device__ __attribute__((always_inline)) inline void split_pair_f32(float v0, float v1, uint32_t (&t)[S]) {
    f32x2 r = {v0, v1};
#pragma unroll
    for (int k = 0; k < S; ++k) {
        t[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
        if (k + 1 < S) r = f32x2{r.x - __uint_as_float(t[k] << 16), r.y - __uint_as_float(t[k] & 0xffff0000u)};
    }
}
template <int S>
__device__ __attribute__((always_inline)) inline void split_f32(float v, uint16_t (&t)[S]) {
    float r = v;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const __bf16 h = static_cast<__bf16>(r);
        t[k] = __builtin_bit_cast(uint16_t, h);
        if (k + 1 < S) r -= static_cast<float>(h);
    }
}

// Coefficients of a folded row (operand code below -1; a partially kept
// pure block's Cartesian component k): sum over the block's kept pure AOs r
// of T[r][k] g[k] C[r][slot], for N consecutive slots from `slot`, in FP64.
// ent and ebits are the screening block's entry tables.
template <int N>
__device__ __attribute__((always_inline)) inline void folded_row(const MolView& m, const MoOperand& mo,
                                                               const int32_t* ent, const uint32_t* ebits, int code,
                                                               int slot, double (&out)[N]) {
    const int c = kFoldedRow - code;
    const int q = c >> 5, k = c & 31;
    const DevShell& sh = m.shells[ent[q] & 0xffffff];
    const uint32_t bits = ebits[q];
    const double* tg = m.tg + sh.l * kMaxPure * kMaxCart + k;
#pragma unroll
    for (int i = 0; i < N; ++i)
        out[i] = 0.0;
    for (int r = 0; r < sh.count; ++r) {
        const double t = tg[r * kMaxCart];
        if (!((bits >> r) & 1u) || t == 0.0) continue;
        const double* src = mo.op64 + static_cast<long long>(sh.prow + r) * mo.slots + slot;
#pragma unroll
        for (int i = 0; i < N; ++i)
            out[i] += t * src[i];
    }
}

// Upper bounds of a folded row's norms over the alpha and the beta slots
// (triangle inequality over its kept pure AOs), for the error scales.
__device__ __attribute__((always_inline)) inline float2 folded_norm(const MolView& m, const MoOperand& mo,
                                                                  const int32_t* ent, const uint32_t* ebits,
                                                                  int code) {
    const int c = kFoldedRow - code;
    const int q = c >> 5, k = c & 31;
    const DevShell& sh = m.shells[ent[q] & 0xffffff];
    const uint32_t bits = ebits[q];
    const double* tg = m.tg + sh.l * kMaxPure * kMaxCart + k;
    float na = 0.f, nb = 0.f;
    for (int r = 0; r < sh.count; ++r) {
        if (!((bits >> r) & 1u)) continue;
        const float t = fabsf(static_cast<float>(tg[r * kMaxCart]));
        const float2 n = mo.norm[sh.prow + r];
        na += t * n.x;
        nb += t * n.y;
    }
    return float2{na, nb};
}

// bf16 term `plane` of x: each term is the bf16 rounding (to nearest even)
// of what the previous terms leave, the residuals exact in FP64 (the split of
// the operand table's planes).
__device__ __attribute__((always_inline)) inline uint16_t bf16_term(double x, int plane) {
    double r = x;
    uint16_t h = 0;
    for (int t = 0; t <= plane; ++t) {
        h = __builtin_bit_cast(uint16_t, static_cast<__bf16>(static_cast<float>(r)));
        r -= static_cast<double>(__uint_as_float(static_cast<uint32_t>(h) << 16));
    }
    return h;
}

// bf16 plane `plane` of 8 FP64 coefficients, as a 16-byte granule of the
// staged table.
__device__ __attribute__((always_inline)) inline uint4 bf16_granule(const double (&f)[8], int plane) {
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        w[i] = static_cast<uint32_t>(bf16_term(f[2 * i], plane)) |
               (static_cast<uint32_t>(bf16_term(f[2 * i + 1], plane)) << 16);
    return uint4{w[0], w[1], w[2], w[3]};
}

// The row of the fold workspace that code (<= kFoldedWs) names, from slot
// `slot` on.
__device__ __attribute__((always_inline)) inline const double* fold_ws_row(const KeptTables& t, int code, int slot) {
    return t.fold_rows + static_cast<long long>(kFoldedWs - code) * t.fold_slots + slot;
}

// bf16 plane `plane` of 8 slots of a fold workspace row (16-byte aligned).
__device__ __attribute__((always_inline)) inline uint4 fold_ws_bf16(const double* p, int plane) {
    double f[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double2 v = reinterpret_cast<const double2*>(p)[i];
        f[2 * i] = v.x;
        f[2 * i + 1] = v.y;
    }
    return bf16_granule(f, plane);
}

// bf16 plane `plane` of a folded row at 8 slots from `slot`, as the 16-byte
// granule of the staged table.
__device__ __attribute__((always_inline)) inline uint4 folded_bf16(const MolView& m, const MoOperand& mo,
                                                                 const int32_t* ent, const uint32_t* ebits, int code,
                                                                 int slot, int plane) {
    double f[8];
    folded_row<8>(m, mo, ent, ebits, code, slot, f);
    return bf16_granule(f, plane);
}

} // namespace aomo
} // namespace kumo
