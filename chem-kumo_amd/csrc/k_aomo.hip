// Fused AO->MO contraction (density_derivatives_mo, src/density.cpp:234-351).
//
// One workgroup (4 waves) owns T points of one screening block and up to 128
// MO slots (alpha then beta occupied MOs, each padded to 16). It walks the
// block's kept AO blocks chunk by chunk (whole blocks, at most 32 Cartesian
// component rows per chunk). The contraction runs over Cartesian components:
// since the pure transform T is linear, C_pure^T (T Phi_cart) =
// (T^T C_pure)^T Phi_cart, so
//   1. every lane evaluates the Cartesian components of one AO block at one
//      point in FP64 (ao_eval.hpp) and writes the needed derivative panels
//      into LDS in the contraction precision (FP64, FP32 or bf16 hi/mid/lo
//      splits);
//   2. the chunk's coefficient rows C_cart = T^T C_kept are formed in FP64
//      from the occupied-MO coefficients of the block's kept AOs only (the
//      screening mask is applied here) and staged in LDS;
//   3. MFMA tiles accumulate Psi[panel] = C_cart^T Phi_cart[panel]
//      (16 MOs x 16 points per tile, all panels of a tile in one wave).
// The epilogue forms the per-point dot products over MOs in FP64 (value,
// gradient, tau terms, Hessian) and reduces them over the tiles of each
// spin channel in a fixed order, so the result is independent of batching.
//
// MFMA operand layouts (wave64): for 16x16xK the lane l holds A row l % 16
// and B column l % 16 at a K slice selected by l / 16; the accumulator lane l
// holds D rows 4 (l / 16) + r, column l % 16. The K slice order is the same
// for A and B, so the sum over K is exact whatever order the hardware uses.
#include "kernels.hpp"

#include "../../include/kumo_hip.h"
#include "common.hpp"

namespace kumo {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

constexpr int KC = kChunkRows; // 32 AO rows per chunk

template <int PREC> struct Prec;
template <> struct Prec<KUMO_HIP_FP64> {
    using E = double;      // LDS element
    using Acc = f64x4;
    static constexpr int splits = 1;
    static constexpr int T = 16;   // points per workgroup
    static constexpr int TPW = 2;  // tiles per wave (8 m-tiles x 1 n-tile / 4 waves)
    static constexpr int LD = 36;  // LDS row stride (elements)
};
template <> struct Prec<KUMO_HIP_FP32> {
    using E = float;
    using Acc = f32x4;
    static constexpr int splits = 1;
    static constexpr int T = 32;
    static constexpr int TPW = 4;
    static constexpr int LD = 36;
};
template <> struct Prec<KUMO_HIP_BF16X3> {
    using E = uint16_t;
    using Acc = f32x4;
    static constexpr int splits = 2;
    static constexpr int T = 32;
    static constexpr int TPW = 4;
    static constexpr int LD = 40;
};
template <> struct Prec<KUMO_HIP_BF16X6> {
    using E = uint16_t;
    using Acc = f32x4;
    static constexpr int splits = 3;
    static constexpr int T = 32;
    static constexpr int TPW = 4;
    static constexpr int LD = 40;
};

__device__ inline uint16_t bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Writes value v (FP64) as the element(s) of split planes.
template <int PREC>
__device__ inline void store_split(typename Prec<PREC>::E* base, size_t plane, double v) {
    if constexpr (PREC == KUMO_HIP_FP64) {
        base[0] = v;
    } else if constexpr (PREC == KUMO_HIP_FP32) {
        base[0] = static_cast<float>(v);
    } else {
        const uint16_t hi = bf16_rn(static_cast<float>(v));
        const double r1 = v - static_cast<double>(bf16_to_f(hi));
        const uint16_t mid = bf16_rn(static_cast<float>(r1));
        base[0] = hi;
        base[plane] = mid;
        if constexpr (Prec<PREC>::splits == 3) {
            const double r2 = r1 - static_cast<double>(bf16_to_f(mid));
            base[2 * plane] = bf16_rn(static_cast<float>(r2));
        }
    }
}

template <int PREC, int ORDER, unsigned PSET>
__global__ void __launch_bounds__(256) aomo_kernel(AoMoLaunch a) {
    using P = Prec<PREC>;
    using E = typename P::E;
    constexpr int T = P::T, NT = T / 16, LD = P::LD, S = P::splits, TPW = P::TPW;
    constexpr int NP = popcount10(PSET);
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    // LDS: phi[S][NP][T][LD], cs[S][128][LD], red[8][kMoments][T] (double)
    E* phi = reinterpret_cast<E*>(smem);
    const size_t phi_plane = static_cast<size_t>(NP) * T * LD;
    E* cs = phi + S * phi_plane;
    const size_t cs_plane = static_cast<size_t>(128) * LD;
    double* red = reinterpret_cast<double*>(cs + S * cs_plane);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Tile: (screening block, point range) and MO slot group.
    const int tiles_per_block = (a.block_size + T - 1) / T;
    const int blk_local = blockIdx.x / tiles_per_block;
    const int sub = blockIdx.x % tiles_per_block;
    const long long blk = a.first_block + blk_local;
    const long long bstart = blk * a.block_size;
    const int bcount = static_cast<int>(min(static_cast<long long>(a.block_size), a.npoints - bstart));
    const int p0 = sub * T;
    if (p0 >= bcount) return; // uniform over the workgroup
    const int nt_valid = min(T, bcount - p0);
    const int slot_lo = a.slot0 + 128 * blockIdx.y;
    const int slot_hi = min(a.slot1, slot_lo + 128);
    const int MT = (slot_hi - slot_lo) / 16;
    const int ntiles = MT * NT;

    const MolView& m = a.mol;
    const int nsh = m.nshells;
    const int32_t* ks = a.tables.kept_shell + static_cast<long long>(blk_local) * nsh;
    const uint32_t* kb = a.tables.kept_bits + static_cast<long long>(blk_local) * nsh;
    const uint8_t* kr = a.tables.kept_row + static_cast<long long>(blk_local) * nsh;
    const int32_t* cb = a.tables.chunk_begin + static_cast<long long>(blk_local) * (nsh + 1);
    const int nchunks = a.tables.nchunks[blk_local];

    typename P::Acc acc[TPW][NP];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int i = 0; i < NP; ++i)
            acc[t][i] = typename P::Acc{0, 0, 0, 0};

    // Point coordinates of this lane's column for AO generation.
    for (int c = 0; c < nchunks; ++c) {
        const int s_first = cb[c], s_last = cb[c + 1];
        const int nsc = s_last - s_first;
        int rows_used = 0;
        {
            const DevShell& lastsh = m.shells[ks[s_last - 1]];
            rows_used = kr[s_last - 1] + lastsh.ncart;
        }
        // 1. Cartesian AO components of the chunk's blocks at the tile's points
        //    (the block's mask and pure transform live in the staged C).
        for (int item = tid; item < nsc * T; item += 256) {
            const int j = item % T;
            const int q = s_first + item / T;
            const DevShell sh = m.shells[ks[q]];
            const int row0 = kr[q];
            E* dst = phi + static_cast<size_t>(j) * LD + row0;
            bool inside = false;
            double x = 0, y = 0, z = 0, r2 = 0;
            if (j < nt_valid) {
                double px, py, pz;
                point_coords(a.ps, bstart + p0 + j, px, py, pz);
                x = px - sh.o[0];
                y = py - sh.o[1];
                z = pz - sh.o[2];
                r2 = sq_norm(x, y, z);
                inside = r2 <= sh.cutoff_r2;
            }
            if (inside) {
                eval_block_cartesian<ORDER, PSET>(sh, m, x, y, z, r2, [&](int k, const double (&v)[NP]) __attribute__((always_inline)) {
#pragma unroll
                    for (int i = 0; i < NP; ++i)
                        store_split<PREC>(dst + static_cast<size_t>(i) * T * LD + k, phi_plane, v[i]);
                });
            } else {
                for (int k = 0; k < sh.ncart; ++k)
#pragma unroll
                    for (int i = 0; i < NP; ++i)
                        store_split<PREC>(dst + static_cast<size_t>(i) * T * LD + k, phi_plane, 0.0);
            }
        }
        // Rows past the chunk's AOs are zero in both operands.
        for (int item = tid; item < (KC - rows_used) * T * NP; item += 256) {
            const int row = rows_used + item % (KC - rows_used);
            const int rest = item / (KC - rows_used);
            const int j = rest % T, i = rest / T;
#pragma unroll
            for (int sp = 0; sp < S; ++sp)
                phi[sp * phi_plane + (static_cast<size_t>(i) * T + j) * LD + row] = E(0);
        }
        // 2. MO coefficients of the chunk's Cartesian rows, cs[slot][row]:
        //    C_cart[k] = sum over kept pure AOs r of T[r][k] C[r] (pure blocks)
        //    or C[k] if AO k is kept (Cartesian blocks), in FP64, then split.
        const int width = slot_hi - slot_lo;
        for (int item = tid; item < KC * width; item += 256) {
            const int slot = item % width;
            const int row = item / width;
            E* d = cs + static_cast<size_t>(slot) * LD + row;
            double v = 0.0;
            if (row < rows_used) {
                int q = s_first;
                while (q + 1 < s_last && kr[q + 1] <= row)
                    ++q;
                const DevShell& sh = m.shells[ks[q]];
                const uint32_t bits = kb[q];
                const int k = row - kr[q];
                const double* c = a.mo.c64 + static_cast<long long>(sh.ao) * a.mo.slots + slot_lo + slot;
                if (!sh.pure) {
                    if ((bits >> k) & 1u) v = c[static_cast<long long>(k) * a.mo.slots];
                } else {
                    const double* t = m.tdense + (sh.l * kMaxPure) * kMaxCart + k;
                    for (int r = 0; r < sh.count; ++r)
                        if ((bits >> r) & 1u) v += t[r * kMaxCart] * c[static_cast<long long>(r) * a.mo.slots];
                }
            }
            store_split<PREC>(d, cs_plane, v);
        }
        __syncthreads();
        // 3. MFMA over the chunk's rows.
        const int g = lane >> 4, r16 = lane & 15;
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int tile = wave + 4 * t;
            if (tile >= ntiles) continue;
            const int mt = tile / NT, nt = tile % NT;
            const E* arow = cs + static_cast<size_t>(mt * 16 + r16) * LD;
            const E* brow = phi + static_cast<size_t>(nt * 16 + r16) * LD;
            if constexpr (PREC == KUMO_HIP_FP64) {
                for (int k0 = 0; k0 < rows_used; k0 += 8) {
                    const double2 av = *reinterpret_cast<const double2*>(arow + k0 + 2 * g);
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        const double2 bv = *reinterpret_cast<const double2*>(brow + static_cast<size_t>(i) * T * LD + k0 + 2 * g);
                        acc[t][i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.y, acc[t][i], 0, 0, 0);
                    }
                }
            } else if constexpr (PREC == KUMO_HIP_FP32) {
                for (int k0 = 0; k0 < rows_used; k0 += 16) {
                    const float4 av = *reinterpret_cast<const float4*>(arow + k0 + 4 * g);
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        const float4 bv = *reinterpret_cast<const float4*>(brow + static_cast<size_t>(i) * T * LD + k0 + 4 * g);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc[t][i], 0, 0, 0);
                    }
                }
            } else {
                bf16x8 av[S];
#pragma unroll
                for (int sp = 0; sp < S; ++sp)
                    av[sp] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const s16x8*>(arow + sp * cs_plane + 8 * g));
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    bf16x8 bv[S];
#pragma unroll
                    for (int sp = 0; sp < S; ++sp)
                        bv[sp] = __builtin_bit_cast(
                            bf16x8, *reinterpret_cast<const s16x8*>(brow + sp * phi_plane + static_cast<size_t>(i) * T * LD + 8 * g));
                    // small terms first
                    if constexpr (S == 3) {
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[2], bv[0], acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1], bv[1], acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0], bv[2], acc[t][i], 0, 0, 0);
                    }
                    acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1], bv[0], acc[t][i], 0, 0, 0);
                    acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0], bv[1], acc[t][i], 0, 0, 0);
                    acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0], bv[0], acc[t][i], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }

    // Epilogue: dot products over the MOs of each tile, FP64.
    // Raw sums q: 0 v0.v0, 1-3 v0.va, 4-6 va.va, 7-12 (v0.vac + va.vc).
    constexpr int kPanelOf[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    (void)kPanelOf;
    const int g = lane >> 4, r16 = lane & 15;
    (void)g;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = wave + 4 * t;
        if (tile >= ntiles) continue;
        const int mt = tile / NT, nt = tile % NT;
        double q[kMoments];
#pragma unroll
        for (int k = 0; k < kMoments; ++k)
            q[k] = 0.0;
        const auto V = [&](int comp, int r) -> double {
            // output panel index of derivative component `comp`
            int idx = -1, n = 0;
#pragma unroll
            for (int c = 0; c < 10; ++c)
                if ((PSET >> c) & 1u) {
                    if (c == comp) idx = n;
                    ++n;
                }
            return idx < 0 ? 0.0 : static_cast<double>(acc[t][idx < 0 ? 0 : idx][r]);
        };
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double v0 = V(0, r);
            q[0] += v0 * v0;
            if constexpr (ORDER >= 1) {
                const double va[3] = {V(1, r), V(2, r), V(3, r)};
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    q[1 + ax] += v0 * va[ax];
                    q[4 + ax] += va[ax] * va[ax];
                }
                if constexpr (ORDER == 2) {
                    constexpr int hc[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
                    int h = 0;
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
                        for (int c2 = ax; c2 < 3; ++c2, ++h)
                            q[7 + h] += v0 * V(4 + hc[ax][c2], r) + va[ax] * va[c2];
                }
            }
        }
        // Sum over the four row groups of the tile.
#pragma unroll
        for (int k = 0; k < kMoments; ++k) {
            q[k] += __shfl_xor(q[k], 16, 64);
            q[k] += __shfl_xor(q[k], 32, 64);
        }
        if (lane < 16)
#pragma unroll
            for (int k = 0; k < kMoments; ++k)
                red[(static_cast<size_t>(mt) * kMoments + k) * T + nt * 16 + r16] = q[k];
    }
    __syncthreads();
    // Sum the m-tiles of each channel in order and store the moments.
    const int alpha_tiles_end = max(0, min(MT, (a.mo.alpha_slots - slot_lo) / 16));
    for (int item = tid; item < 2 * kMoments * T; item += 256) {
        const int j = item % T;
        const int k = (item / T) % kMoments;
        const int ch = item / (T * kMoments);
        if (j >= nt_valid) continue;
        const int mt0 = ch == 0 ? 0 : alpha_tiles_end, mt1 = ch == 0 ? alpha_tiles_end : MT;
        double s = 0.0;
        for (int mt = mt0; mt < mt1; ++mt)
            s += red[(static_cast<size_t>(mt) * kMoments + k) * T + j];
        const double f = k == 0 ? 1.0 : (k <= 3 ? 2.0 : (k <= 6 ? 0.5 : 2.0));
        const long long point = bstart + p0 + j - a.batch_first;
        double* dst = a.moments + (static_cast<long long>(blockIdx.y) * 2 + ch) * kMoments * a.batch_points;
        dst[k * a.batch_points + point] = f * s;
    }
}

template <int PREC, int ORDER, unsigned PSET> void launch_one(const AoMoLaunch& a, hipStream_t stream) {
    using P = Prec<PREC>;
    constexpr int NP = popcount10(PSET);
    const size_t lds = sizeof(typename P::E) * P::splits * (static_cast<size_t>(NP) * P::T * P::LD + 128 * P::LD) +
                       sizeof(double) * 8 * kMoments * P::T;
    const int tiles_per_block = (a.block_size + P::T - 1) / P::T;
    const int groups = (a.slot1 - a.slot0 + 127) / 128;
    const dim3 grid(static_cast<unsigned>(a.nblk * tiles_per_block), static_cast<unsigned>(groups));
    static bool configured = false;
    if (!configured) {
        KUMO_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&aomo_kernel<PREC, ORDER, PSET>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        configured = true;
    }
    aomo_kernel<PREC, ORDER, PSET><<<grid, dim3(256), lds, stream>>>(a);
    KUMO_HIP(hipGetLastError());
}

template <int PREC> void launch_prec(const AoMoLaunch& a, hipStream_t stream) {
    if (a.order == 0) launch_one<PREC, 0, kPanelsValue>(a, stream);
    else if (a.order == 1) launch_one<PREC, 1, kPanelsGrad>(a, stream);
    else if (a.hessian == 2) launch_one<PREC, 2, kPanelsAll>(a, stream);
    else launch_one<PREC, 2, kPanelsDiag>(a, stream);
}

} // namespace

void launch_ao_mo(const AoMoLaunch& a, hipStream_t stream) {
    if (a.nblk <= 0 || a.slot1 <= a.slot0) return;
    switch (a.precision) {
    case KUMO_HIP_FP64: launch_prec<KUMO_HIP_FP64>(a, stream); break;
    case KUMO_HIP_FP32: launch_prec<KUMO_HIP_FP32>(a, stream); break;
    case KUMO_HIP_BF16X3: launch_prec<KUMO_HIP_BF16X3>(a, stream); break;
    default: launch_prec<KUMO_HIP_BF16X6>(a, stream); break;
    }
}

} // namespace kumo
