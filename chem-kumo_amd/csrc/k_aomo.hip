// Fused AO->MO contraction (density_derivatives_mo, src/density.cpp:234-351).
//
// One workgroup (4 waves) owns T points of one screening block and up to 128
// MO slots (alpha then beta occupied MOs, each padded to 16). It walks the
// block's contraction chunks (at most 32 operand rows each, built by the mask
// kernel from the kept AO blocks, ordered by descending l). Per chunk:
//   1. the chunk's rows of the MO operand table are copied to LDS with
//      coalesced 16-byte loads: for a fully kept block these are its
//      Cartesian rows g[k] (T^T C)[k] (pure transform and component factor
//      folded into the coefficients), for a partially kept pure block its
//      plain pure rows C[r]; rows of dropped AOs are zero (the screening mask);
//   2. each thread evaluates 4 (FP64: 2) consecutive Phi rows of the chunk
//      at one point in FP64 from the rows' AO block radial sums: unscaled
//      Cartesian components P_k R for Cartesian rows, or the pure AOs
//      sum_k T g P_k R for pure rows, converts the needed derivative panels to
//      the contraction precision (FP64, FP32 or bf16 splits) and writes them
//      with one vector store per plane and panel;
//   3. MFMA tiles accumulate Psi[panel] = C_rows^T Phi_rows[panel]
//      (16 MOs x 16 points per tile, all panels of a tile in one wave; a wave
//      keeps one 16-point column of tiles, so each Phi operand it loads feeds
//      all of its MO tiles).
// The epilogue forms the per-point dot products over MOs in FP64 (value,
// gradient, tau terms, Hessian) and reduces them over the tiles of each
// spin channel in a fixed order, so the result is independent of batching.
// Reduced precisions also form, per point and channel, the 2-norm over the
// operand rows k of Phi_k times the norm of coefficient row k over the
// channel's MOs, and (with gradient panels) of |grad Phi_k| times the same
// norm: the scales of the contraction's rounding error that the rescue tests.
//
// MFMA operand layouts (wave64): for 16x16xK the lane l holds A row l % 16
// and B column l % 16 at a K slice selected by l / 16; accumulator register r
// of lane l holds D column l % 16 and row 4 (l / 16) + r for the FP32
// accumulators, row l / 16 + 4 r for the FP64 one (its vector width is 1).
// The K slice order is the same for A and B, so the sum over K is exact
// whatever order the hardware uses.
// LDS strides are chosen so that the 16 lanes of each K slice hit distinct
// banks. FP32 and FP64: coefficient rows have a pitch of 132 and 144
// elements, the Phi row of a point 36 and 34. bf16: coefficient rows are
// staged in pairs, one 32-bit word (row 2q low, row 2q + 1 high) per slot,
// 132 words per pair, so a lane reads its 8 K-consecutive values as four
// words; the 32 Phi values of a point are four 8-row granules whose order
// is XOR-swizzled by (point / 4) % 4, so 16 consecutive points read
// distinct banks without padding.
//
// Split-bf16 precisions: both operands are sums of bf16 terms (Phi split in
// the kernel, the coefficients when the molecule is loaded) and the MFMA
// products of the terms whose combined order is within the precision are
// accumulated in FP32: BF16X3 two terms, 3 products; BF16X6 three terms,
// 6 products. More terms gain nothing: the bf16 MFMA aligns the products of
// each 8-element K group to the group's largest and keeps about 24 bits, so
// the contraction error is about 2^-24 sum_k |C_k Phi_k| whatever the split.
#include "kernels.hpp"

#include "../../include/kumo_hip.h"
#include "common.hpp"
#include "aomo_common.hpp"

namespace kumo {
namespace {
using namespace aomo;

constexpr int KC = kChunkRows; // 32 operand rows per chunk

template <int PREC> struct Prec;
template <> struct Prec<KUMO_HIP_FP64> {
    using E = double;
    using G = double; // arithmetic of the Phi components
    using Acc = f64x4;
    static constexpr int S = 1;       // split planes
    static constexpr int T = 16;      // points per workgroup
    static constexpr int LDK = 34;    // Phi elements per point row
    static constexpr int PITCH = 144; // coefficient row pitch (elements)
};
template <> struct Prec<KUMO_HIP_FP32> {
    using E = float;
    using G = double;
    using Acc = f32x4;
    static constexpr int S = 1;
    static constexpr int LDK = 36;
    static constexpr int PITCH = 132;
};
template <int NSPLIT> struct PrecBf16 {
    using E = uint16_t;
    using G = double;
    using Acc = f32x4;
    static constexpr int S = NSPLIT;
    static constexpr int LDK = 32;    // swizzled 8-row granules
    static constexpr int PITCH = 132; // 32-bit words per coefficient row pair
};
template <> struct Prec<KUMO_HIP_BF16X3> : PrecBf16<2> {};
template <> struct Prec<KUMO_HIP_BF16X6> : PrecBf16<3> {};

template <int PREC> constexpr bool is_bf16 = PREC == KUMO_HIP_BF16X3 || PREC == KUMO_HIP_BF16X6;

// Element offset of row r (< 32) of point j in a bf16 Phi plane.
__device__ __attribute__((always_inline)) inline int phi_bf16_offset(int j, int r) {
    return j * 32 + ((((r >> 3) ^ (j >> 2)) & 3) << 3) + (r & 7);
}

// Workgroups per CU the register budget is sized for: 32-point tiles keep
// 4 x NP accumulator tiles per wave, reduced-precision 16-point tiles half as
// many and run three workgroups per CU (FP64 tiles are always 16 points).
template <int PREC, unsigned PSET, int TP> constexpr int occupancy() {
    return popcount10(PSET) > 7 ? 1 : (TP == 16 && PREC != KUMO_HIP_FP64 ? 3 : 2);
}

// One workgroup tile: T points of a screening block (tile index tx in the
// grid's x order) or, in rescue mode, rescue tile tx.
// WSUM: weighted mode (orbital-energy sums, FP64 values only): instead of the
// channel moments, per point the sums over slots i of w_s(i) psi_i^2.
template <int PREC, int ORDER, unsigned PSET, int TP, bool WSUM>
__device__ __attribute__((always_inline)) inline void aomo_tile(const AoMoLaunch& a, const int tx) {
    using P = Prec<PREC>;
    using E = typename P::E;
    static_assert(TP == 16 || TP == 32, "16 or 32 points per workgroup");
    constexpr int T = TP, NT = T / 16, S = P::S, LDK = P::LDK, PITCH = P::PITCH;
    constexpr int NP = popcount10(PSET);
    constexpr int TPW = 8 * NT / 4; // tiles per wave at 128 slots
    constexpr int PHI_PLANE = NP * T * LDK;
    constexpr int CS_PLANE = KC * PITCH;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    // LDS: phi[S][NP][T][LDK], cs[S][KC][PITCH]; after the chunk loop the
    // same memory holds red[MT][kMoments][T] (double), then esl[256][4]
    // (float, the threads' error-scale sums: value alpha, beta, gradient
    // alpha, beta).
    E* phi = reinterpret_cast<E*>(smem);
    E* cs = phi + S * PHI_PLANE;
    double* red = reinterpret_cast<double*>(smem);
    float* esl = reinterpret_cast<float*>(smem + sizeof(double) * 8 * kMoments * T);
    const bool scale = PREC != KUMO_HIP_FP64 && a.rho_scale != nullptr; // uniform
    float es0 = 0.f, es1 = 0.f, eg0 = 0.f, eg1 = 0.f;
    constexpr bool kGradScale = ORDER >= 1 && (PSET & 0xfu) == 0xfu; // panels 1-3 are the gradient

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Tile: T points of one screening block, consecutive or (rescue mode)
    // listed by their offsets in the block.
    int blk_local, p0 = 0, nt_valid;
    const uint16_t* roff = nullptr;
    if (a.r_tile_blk) {
        const int t = tx;
        blk_local = a.r_tile_blk[t];
        const int sub = a.r_tile_sub[t];
        nt_valid = min(T, a.r_count[blk_local] - sub * T);
        roff = a.r_offs + static_cast<long long>(blk_local) * a.block_size + sub * T;
    } else {
        const int tiles_per_block = (a.block_size + T - 1) / T;
        blk_local = tx / tiles_per_block;
        p0 = (tx % tiles_per_block) * T;
        const long long bs0 = (a.first_block + blk_local) * a.block_size;
        const int bcount = static_cast<int>(min(static_cast<long long>(a.block_size), a.npoints - bs0));
        if (p0 >= bcount) return; // uniform over the workgroup
        nt_valid = min(T, bcount - p0);
    }
    const long long blk = a.first_block + blk_local;
    const long long bstart = blk * a.block_size;
    const int slot_lo = a.slot0 + 128 * blockIdx.y;
    const int width = min(a.slot1, slot_lo + 128) - slot_lo;
    const int MT = width / 16;
    const int ntiles = MT * NT;

    const MolView& m = a.mol;
    const int nsh = m.nshells;
    const int32_t* ent = a.tables.ent + static_cast<long long>(blk_local) * nsh;
    const uint32_t* ebits = a.tables.ent_bits + static_cast<long long>(blk_local) * nsh;
    const uint8_t* erow = a.tables.ent_row + static_cast<long long>(blk_local) * nsh;
    const int32_t* cb = a.tables.chunk_begin + static_cast<long long>(blk_local) * (a.tables.max_chunks + 1);
    const int32_t* crows = a.tables.chunk_rows + static_cast<long long>(blk_local) * a.tables.max_chunks * KC * 2;
    const int nchunks = a.tables.nchunks[blk_local];

    // This thread's point for AO generation (T divides 256).
    const int j = tid % T;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (j < nt_valid) point_coords(a.ps, bstart + (roff ? roff[j] : p0 + j), px, py, pz);

    typename P::Acc acc[TPW][NP];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int i = 0; i < NP; ++i)
            acc[t][i] = typename P::Acc{0, 0, 0, 0};

    const E* op;
    if constexpr (PREC == KUMO_HIP_FP64) op = reinterpret_cast<const E*>(a.mo.op64);
    else if constexpr (PREC == KUMO_HIP_FP32) op = reinterpret_cast<const E*>(a.mo.op32);
    else op = reinterpret_cast<const E*>(a.mo.opbf);
    const long long op_plane = static_cast<long long>(a.mo.rows) * a.mo.slots;
    constexpr int EPG = 16 / static_cast<int>(sizeof(E)); // elements per 16-byte granule
    constexpr int GMAX = 128 / EPG;                        // granules of a 128-slot row (power of two)
    const int G = width / EPG;                             // granules per coefficient row

    // The chunk tables a thread reads (operand rows of its staging items, its
    // Phi rows' operand rows and row descriptors, the chunk's used rows) are
    // fetched one chunk ahead, before the MFMA phase, so that their load
    // latency is not on the next chunk's path.
    constexpr int NSTAGE = is_bf16<PREC> ? S * (KC / 2) * GMAX / 256 : S * KC * GMAX / 256;
    constexpr int RPL = KC * T / 256;
    const int r0 = (tid / T) * RPL;
    int pg0[NSTAGE], pg1[NSTAGE], prow_g[RPL], prow_info[RPL], next_rows_used = 0;
    const auto fetch = [&](int cc) __attribute__((always_inline)) {
        const int32_t* crc = crows + 2 * cc * KC;
#pragma unroll
        for (int u = 0; u < NSTAGE; ++u) {
            const int item = tid + 256 * u;
            if constexpr (is_bf16<PREC>) {
                const int q = (item / GMAX) % (KC / 2);
                pg0[u] = crc[4 * q];
                pg1[u] = crc[4 * q + 2];
            } else {
                pg0[u] = crc[2 * ((item / GMAX) % KC)];
                pg1[u] = -1;
            }
        }
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            prow_g[i] = crc[2 * (r0 + i)];
            prow_info[i] = crc[2 * (r0 + i) + 1];
        }
        const int e1n = cb[cc + 1];
        const DevShell& ls = m.shells[ent[e1n - 1] & 0xffffff];
        next_rows_used = erow[e1n - 1] + ls.ncart;
    };
    if (nchunks > 0) fetch(0);

    for (int c = 0; c < nchunks; ++c) {
        const int rows_used = next_rows_used;
        int g_row[RPL], info_row[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            g_row[i] = prow_g[i];
            info_row[i] = prow_info[i];
        }
        // 1. Coefficient rows of the chunk (zero past its rows). Folded rows
        // (partially kept pure blocks, rare) come from the fold workspace, or
        // beyond it are formed in a second pass, so that the plain copies keep
        // their loads together.
        bool deferred = false;
        const auto entry_fold = [](int g) { return g < -1 && g > kFoldedWs; };
        if constexpr (is_bf16<PREC>) {
            // rows 2q and 2q + 1, 8 slots each, interleaved into 32-bit words
            const auto item_bf16 = [&](int u, bool fold) __attribute__((always_inline)) {
                const int item = tid + 256 * u;
                const int gi = item % GMAX;
                const int q = (item / GMAX) % (KC / 2);
                const int sp = item / ((KC / 2) * GMAX);
                if (gi >= G) return;
                const int g0 = pg0[u], g1 = pg1[u];
                const E* src = op + sp * op_plane + slot_lo + gi * EPG;
                uint4 v0 = uint4{0, 0, 0, 0}, v1 = uint4{0, 0, 0, 0};
                const bool late = entry_fold(g0) || entry_fold(g1);
                if (late != fold) {
                    deferred = deferred || late;
                    return;
                }
                if (!fold) {
                    // operand table rows and fold workspace rows: one load each
                    const E* ws = static_cast<const E*>(a.tables.fold_lo) +
                                  static_cast<long long>(sp) * a.tables.fold_cap * a.tables.fold_slots + slot_lo +
                                  gi * EPG;
                    const auto row = [&](int g) __attribute__((always_inline)) {
                        return g >= 0 ? src + static_cast<long long>(g) * a.mo.slots
                                      : ws + static_cast<long long>(kFoldedWs - g) * a.tables.fold_slots;
                    };
                    if (g0 >= 0 || g0 <= kFoldedWs) v0 = *reinterpret_cast<const uint4*>(row(g0));
                    if (g1 >= 0 || g1 <= kFoldedWs) v1 = *reinterpret_cast<const uint4*>(row(g1));
                } else {
                    const auto granule = [&](int g) __attribute__((always_inline)) {
                        if (g >= 0) return *reinterpret_cast<const uint4*>(src + static_cast<long long>(g) * a.mo.slots);
                        if (g <= kFoldedWs)
                            return fold_ws_bf16(fold_ws_row(a.tables, g, slot_lo + gi * EPG), sp);
                        if (g < -1) return folded_bf16(m, a.mo, ent, ebits, g, slot_lo + gi * EPG, sp);
                        return uint4{0, 0, 0, 0};
                    };
                    v0 = granule(g0);
                    v1 = granule(g1);
                }
                const auto lo = [](uint32_t x, uint32_t y) { return (x & 0xffffu) | (y << 16); };
                const auto hi = [](uint32_t x, uint32_t y) { return (x >> 16) | (y & 0xffff0000u); };
                uint4* d = reinterpret_cast<uint4*>(reinterpret_cast<uint32_t*>(cs + sp * CS_PLANE) + q * PITCH +
                                                    gi * EPG);
                d[0] = uint4{lo(v0.x, v1.x), hi(v0.x, v1.x), lo(v0.y, v1.y), hi(v0.y, v1.y)};
                d[1] = uint4{lo(v0.z, v1.z), hi(v0.z, v1.z), lo(v0.w, v1.w), hi(v0.w, v1.w)};
            };
#pragma unroll
            for (int u = 0; u < NSTAGE; ++u)
                item_bf16(u, false);
            if (deferred)
                for (int u = 0; u < NSTAGE; ++u)
                    item_bf16(u, true);
        } else {
            const auto item_plain = [&](int u, bool fold) __attribute__((always_inline)) {
                const int item = tid + 256 * u;
                const int gi = item % GMAX;
                const int r = (item / GMAX) % KC;
                const int sp = item / (KC * GMAX);
                if (gi >= G) return;
                const int grow = pg0[u];
                if (entry_fold(grow) != fold) {
                    deferred = deferred || entry_fold(grow);
                    return;
                }
                uint4 v = uint4{0, 0, 0, 0};
                if (grow >= 0) {
                    v = *reinterpret_cast<const uint4*>(op + sp * op_plane + static_cast<long long>(grow) * a.mo.slots +
                                                        slot_lo + gi * EPG);
                } else if (grow <= kFoldedWs) {
                    // the fold workspace as staged (FP64 rows for FP64, FP32 rows for FP32)
                    const long long at = static_cast<long long>(kFoldedWs - grow) * a.tables.fold_slots + slot_lo +
                                         gi * EPG;
                    if constexpr (PREC == KUMO_HIP_FP64) v = *reinterpret_cast<const uint4*>(a.tables.fold_rows + at);
                    else v = *reinterpret_cast<const uint4*>(static_cast<const float*>(a.tables.fold_lo) + at);
                } else if (grow < -1) {
                    double f[EPG];
                    folded_row<EPG>(m, a.mo, ent, ebits, grow, slot_lo + gi * EPG, f);
                    if constexpr (PREC == KUMO_HIP_FP64) {
                        v = __builtin_bit_cast(uint4, double2{f[0], f[1]});
                    } else {
                        float4 h;
                        h.x = static_cast<float>(f[0]);
                        h.y = static_cast<float>(f[1 % EPG]);
                        h.z = static_cast<float>(f[2 % EPG]);
                        h.w = static_cast<float>(f[3 % EPG]);
                        v = __builtin_bit_cast(uint4, h);
                    }
                }
                *reinterpret_cast<uint4*>(cs + sp * CS_PLANE + r * PITCH + gi * EPG) = v;
            };
#pragma unroll
            for (int u = 0; u < NSTAGE; ++u)
                item_plain(u, false);
            if (deferred)
                for (int u = 0; u < NSTAGE; ++u)
                    item_plain(u, true);
        }
        // 2. Phi: thread (point j, row group) evaluates RPL consecutive rows of
        //    the chunk at its point (rows past the chunk's end are zero), in
        //    GT arithmetic (FP64: forming the components in FP32 would double
        //    the reduced precisions' error and the FP64 rescue they need).
        {
            using GT = typename P::G;
            int cached = -1;
            bool inside = false;
            GT R[3] = {GT(0), GT(0), GT(0)}, x = GT(0), y = GT(0), z = GT(0);
            const auto eval_row = [&](int i, GT (&v)[NP]) __attribute__((always_inline)) {
#pragma unroll
                for (int p = 0; p < NP; ++p)
                    v[p] = GT(0);
                const int info = info_row[i];
                if (info < 0 || j >= nt_valid) return;
                const int sidx = info >> 15;
                const ShellRec& rec = m.rec[sidx];
                if (sidx != cached) {
                    cached = sidx;
                    const double dx = px - rec.o[0], dy = py - rec.o[1], dz = pz - rec.o[2];
                    const double r2 = sq_norm(dx, dy, dz);
                    inside = r2 <= rec.cutoff_r2;
                    if (inside) {
                        double R64[3];
                        radial_sums_rec<ORDER>(rec, m, r2, R64);
#pragma unroll
                        for (int q = 0; q < 3; ++q)
                            R[q] = static_cast<GT>(R64[q]);
                    }
                    x = static_cast<GT>(dx);
                    y = static_cast<GT>(dy);
                    z = static_cast<GT>(dz);
                }
                if (!inside) return;
                const int qq[3] = {(info >> 6) & 7, (info >> 3) & 7, info & 7};
                cart_component_pw<ORDER, PSET, GT>(qq, x, y, z, R, v);
            };
            const auto add_scale = [&](int i, const GT (&v)[NP]) __attribute__((always_inline)) {
                if (!scale) return;
                const int g = g_row[i];
                if (g == -1) return;
                const float2 w = g >= 0 ? a.mo.norm[g]
                                        : (g <= kFoldedWs ? a.tables.fold_norm[kFoldedWs - g]
                                                          : folded_norm(m, a.mo, ent, ebits, g));
                const float av = static_cast<float>(v[0]);
                const float wx2 = w.x * w.x, wy2 = w.y * w.y;
                es0 += (av * av) * wx2;
                es1 += (av * av) * wy2;
                if constexpr (kGradScale) {
                    const float gx = static_cast<float>(v[1]), gy = static_cast<float>(v[2]),
                                gz = static_cast<float>(v[3]);
                    const float ag2 = gx * gx + gy * gy + gz * gz;
                    eg0 += ag2 * wx2;
                    eg1 += ag2 * wy2;
                }
            };
            constexpr int PLANE = NP * T * LDK;
            E* d0 = phi + j * LDK + r0;
            if constexpr (is_bf16<PREC>) {
                static_assert(RPL % 2 == 0 && KC == 32 && LDK == 32, "bf16 row pairs, swizzled 32-row points");
                E* db = phi + phi_bf16_offset(j, r0);
#pragma unroll
                for (int pi = 0; pi < RPL / 2; ++pi) {
                    GT va[NP], vb[NP];
                    eval_row(2 * pi, va);
                    eval_row(2 * pi + 1, vb);
                    add_scale(2 * pi, va);
                    add_scale(2 * pi + 1, vb);
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        uint32_t tt[S];
                        if constexpr (sizeof(GT) == 4) split_pair_f32<S>(va[p], vb[p], tt);
                        else split_pair<S>(va[p], vb[p], tt);
                        E* d = db + p * T * LDK + 2 * pi;
#pragma unroll
                        for (int k = 0; k < S; ++k)
                            *reinterpret_cast<uint32_t*>(d + k * PLANE) = tt[k];
                    }
                }
            } else if constexpr (PREC == KUMO_HIP_FP32) {
                static_assert(RPL == 4 || RPL == 2, "two or four FP32 rows per lane");
                float st[RPL][NP];
#pragma unroll
                for (int i = 0; i < RPL; ++i) {
                    GT v[NP];
                    eval_row(i, v);
                    add_scale(i, v);
#pragma unroll
                    for (int p = 0; p < NP; ++p)
                        st[i][p] = static_cast<float>(v[p]);
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    if constexpr (RPL == 4)
                        *reinterpret_cast<float4*>(d0 + p * T * LDK) = float4{st[0][p], st[1][p], st[2][p], st[3][p]};
                    else
                        *reinterpret_cast<float2*>(d0 + p * T * LDK) = float2{st[0][p], st[1][p]};
                }
            } else {
                static_assert(RPL == 2, "two FP64 rows per lane");
                GT va[NP], vb[NP];
                eval_row(0, va);
                eval_row(1, vb);
#pragma unroll
                for (int p = 0; p < NP; ++p)
                    *reinterpret_cast<double2*>(d0 + p * T * LDK) = double2{va[p], vb[p]};
            }
        }
        if (c + 1 < nchunks) fetch(c + 1); // in flight during the MFMA phase
        __syncthreads();
        // 3. MFMA over the chunk's rows. The wave's tiles share one 16-point
        //    column nt and take MO tiles mt = (wave + 4 t) / NT.
        const int g = lane >> 4, r16 = lane & 15;
        const int nt = wave % NT;
        const int jb = nt * 16 + r16; // B column (point) of this lane
        if constexpr (PREC == KUMO_HIP_FP64) {
            for (int k0 = 0; k0 < rows_used; k0 += 4) {
                double av[TPW];
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    const int mt = (wave + 4 * t) / NT;
                    av[t] = cs[(k0 + g) * PITCH + mt * 16 + r16];
                }
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const double bv = phi[i * (T * LDK) + jb * LDK + k0 + g];
#pragma unroll
                    for (int t = 0; t < TPW; ++t)
                        if (wave + 4 * t < ntiles)
                            acc[t][i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t], bv, acc[t][i], 0, 0, 0);
                }
            }
        } else if constexpr (PREC == KUMO_HIP_FP32) {
            for (int k0 = 0; k0 < rows_used; k0 += 16) {
                float av[TPW][4];
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    const int mt = (wave + 4 * t) / NT;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        av[t][e] = cs[(k0 + 4 * g + e) * PITCH + mt * 16 + r16];
                }
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const float4 bv = *reinterpret_cast<const float4*>(phi + i * (T * LDK) + jb * LDK + k0 + 4 * g);
#pragma unroll
                    for (int t = 0; t < TPW; ++t)
                        if (wave + 4 * t < ntiles) {
                            acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t][0], bv.x, acc[t][i], 0, 0, 0);
                            acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t][1], bv.y, acc[t][i], 0, 0, 0);
                            acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t][2], bv.z, acc[t][i], 0, 0, 0);
                            acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t][3], bv.w, acc[t][i], 0, 0, 0);
                        }
                }
            }
        } else {
            bf16x8 av[TPW][S];
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int mt = (wave + 4 * t) / NT;
#pragma unroll
                for (int sp = 0; sp < S; ++sp) {
                    const uint32_t* w = reinterpret_cast<const uint32_t*>(cs + sp * CS_PLANE) + 4 * g * PITCH +
                                        mt * 16 + r16;
                    av[t][sp] = __builtin_bit_cast(bf16x8, uint4{w[0], w[PITCH], w[2 * PITCH], w[3 * PITCH]});
                }
            }
            const int bofs = phi_bf16_offset(jb, 8 * g);
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                bf16x8 bv[S];
#pragma unroll
                for (int sp = 0; sp < S; ++sp)
                    bv[sp] = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const s16x8*>(phi + sp * PHI_PLANE + i * T * LDK + bofs));
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    if (wave + 4 * t >= ntiles) continue;
                    // small terms first
                    if constexpr (S == 3) {
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t][2], bv[0], acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t][1], bv[1], acc[t][i], 0, 0, 0);
                        acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t][0], bv[2], acc[t][i], 0, 0, 0);
                    }
                    acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t][1], bv[0], acc[t][i], 0, 0, 0);
                    acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t][0], bv[1], acc[t][i], 0, 0, 0);
                    acc[t][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t][0], bv[0], acc[t][i], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }

    __syncthreads(); // (a block without kept AOs skips the loop; red aliases Phi)
    if constexpr (WSUM) {
        static_assert(ORDER == 0 && PREC == KUMO_HIP_FP64, "weighted sums of FP64 MO values");
        // Tile rows are MO slots: FP64 accumulator register r of lane l holds
        // slot slot_lo + 16 mt + l / 16 + 4 r at point column l % 16.
        const int r16 = lane & 15;
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int tile = wave + 4 * t;
            if (tile >= ntiles) continue;
            const int mt = tile / NT, nt = tile % NT;
            double q[kOeSums];
#pragma unroll
            for (int k = 0; k < kOeSums; ++k)
                q[k] = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double v = static_cast<double>(acc[t][0][r]);
                const double sq = v * v;
                const double* w = a.weights + slot_lo + mt * 16 + (lane >> 4) + 4 * r;
#pragma unroll
                for (int k = 0; k < kOeSums; ++k)
                    q[k] += w[static_cast<long long>(k) * a.mo.slots] * sq;
            }
#pragma unroll
            for (int k = 0; k < kOeSums; ++k) {
                q[k] += __shfl_xor(q[k], 16, 64);
                q[k] += __shfl_xor(q[k], 32, 64);
            }
            if (lane < 16)
#pragma unroll
                for (int k = 0; k < kOeSums; ++k)
                    red[(static_cast<size_t>(mt) * kOeSums + k) * T + nt * 16 + r16] = q[k];
        }
        __syncthreads();
        // Sum the m-tiles in order and store the sums of this slot group.
        for (int item = tid; item < kOeSums * T; item += 256) {
            const int jj = item % T, k = item / T;
            if (jj >= nt_valid) continue;
            double sum = 0.0;
            for (int mt = 0; mt < MT; ++mt)
                sum += red[(static_cast<size_t>(mt) * kOeSums + k) * T + jj];
            const long long point = bstart + p0 + jj - a.batch_first;
            a.moments[(static_cast<long long>(blockIdx.y) * kOeSums + k) * a.batch_points + point] = sum;
        }
        return;
    }
    if (scale) {
        esl[4 * tid] = es0;
        esl[4 * tid + 1] = es1;
        esl[4 * tid + 2] = eg0;
        esl[4 * tid + 3] = eg1;
    }
    // Epilogue: dot products over the MOs of each tile, FP64.
    // Raw sums q: 0 v0.v0, 1-3 v0.va, 4-6 va.va, 7-12 (v0.vac + va.vc).
    const int r16 = lane & 15;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = wave + 4 * t;
        if (tile >= ntiles) continue;
        const int mt = tile / NT, nt = tile % NT;
        double q[kMoments];
#pragma unroll
        for (int k = 0; k < kMoments; ++k)
            q[k] = 0.0;
        const auto V = [&](int comp, int r) __attribute__((always_inline)) -> double {
            // output panel index of derivative component `comp`
            int idx = -1, n = 0;
#pragma unroll
            for (int cc = 0; cc <= kPanelLap; ++cc)
                if ((PSET >> cc) & 1u) {
                    if (cc == comp) idx = n;
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
                if constexpr (ORDER == 2 && ((PSET >> kPanelLap) & 1u)) {
                    // Laplacian panel: the whole trace in the xx moment
                    q[7] += v0 * V(kPanelLap, r) + va[0] * va[0] + va[1] * va[1] + va[2] * va[2];
                } else if constexpr (ORDER == 2) {
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
            if (!moment_used<ORDER, PSET>(k)) continue;
            q[k] += __shfl_xor(q[k], 16, 64);
            q[k] += __shfl_xor(q[k], 32, 64);
        }
        if (lane < 16)
#pragma unroll
            for (int k = 0; k < kMoments; ++k)
                if (moment_used<ORDER, PSET>(k))
                    red[(static_cast<size_t>(mt) * kMoments + k) * T + nt * 16 + r16] = q[k];
    }
    __syncthreads();
    // Sum the m-tiles of each channel in order and store the moments.
    const int alpha_tiles_end = max(0, min(MT, (a.mo.alpha_slots - slot_lo) / 16));
    for (int item = tid; item < 2 * kMoments * T; item += 256) {
        const int jj = item % T;
        const int k = (item / T) % kMoments;
        const int ch = item / (T * kMoments);
        if (jj >= nt_valid) continue;
        const int mt0 = ch == 0 ? 0 : alpha_tiles_end, mt1 = ch == 0 ? alpha_tiles_end : MT;
        double s = 0.0;
        if (moment_used<ORDER, PSET>(k))
            for (int mt = mt0; mt < mt1; ++mt)
                s += red[(static_cast<size_t>(mt) * kMoments + k) * T + jj];
        const double f = k == 0 ? 1.0 : (k <= 3 ? 2.0 : (k <= 6 ? 0.5 : 2.0));
        const long long point = roff ? static_cast<long long>(tx) * T + jj : bstart + p0 + jj - a.batch_first;
        double* dst = a.moments + (static_cast<long long>(blockIdx.y) * 2 + ch) * kMoments * a.batch_points;
        dst[k * a.batch_points + point] = f * s;
        if (k == 0 && scale && blockIdx.y == 0) {
            float e = 0.f, eg = 0.f;
            for (int q = jj; q < 256; q += T) {
                e += esl[4 * q + ch];
                eg += esl[4 * q + 2 + ch];
            }
            a.rho_scale[ch * a.batch_points + point] = sqrtf(e);
            if (kGradScale) a.rho_scale[(2 + ch) * a.batch_points + point] = sqrtf(eg);
        }
    }
}

// Rescue mode (FP64 only) runs a bounded grid whose workgroups take the
// rescue tiles in turn, so that the grid need not match the tile count known
// only on the device.
template <int PREC, int ORDER, unsigned PSET, int TP, bool WSUM = false>
__global__ void __launch_bounds__(256, (occupancy<PREC, PSET, TP>())) aomo_kernel(AoMoLaunch a) {
    if constexpr (PREC == KUMO_HIP_FP64 && !WSUM) {
        if (a.r_tile_blk) {
            const int total = *a.r_total; // uniform over the workgroup
            for (int t = blockIdx.x; t < total; t += gridDim.x) {
                aomo_tile<PREC, ORDER, PSET, TP, WSUM>(a, t);
                __syncthreads(); // the next tile reuses the LDS
            }
            return;
        }
    }
    aomo_tile<PREC, ORDER, PSET, TP, WSUM>(a, blockIdx.x);
}

// Workgroups of a rescue launch: enough to fill every CU several times.
constexpr unsigned kRescueGroups = 2048;

template <int PREC, int ORDER, unsigned PSET, int TP, bool WSUM = false>
void launch_one(const AoMoLaunch& a, hipStream_t stream) {
    using P = Prec<PREC>;
    constexpr int NP = popcount10(PSET);
    const size_t work = sizeof(typename P::E) * P::S * (static_cast<size_t>(NP) * TP * P::LDK + KC * P::PITCH);
    const size_t red = sizeof(double) * 8 * kMoments * TP + sizeof(float) * 4 * 256;
    const size_t lds = work > red ? work : red;
    const int tiles_per_block = (a.block_size + TP - 1) / TP;
    const int groups = (a.slot1 - a.slot0 + 127) / 128;
    // (rescue mode: at most every block's points in tiles of T, taken in turn
    // by at most kRescueGroups workgroups)
    unsigned gx = static_cast<unsigned>(a.nblk * tiles_per_block);
    if (a.r_tile_blk && gx > kRescueGroups) gx = kRescueGroups;
    const dim3 grid(gx, static_cast<unsigned>(groups));
    static bool configured = false;
    if (!configured) {
        KUMO_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&aomo_kernel<PREC, ORDER, PSET, TP, WSUM>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        configured = true;
    }
    aomo_kernel<PREC, ORDER, PSET, TP, WSUM><<<grid, dim3(256), lds, stream>>>(a);
    KUMO_HIP(hipGetLastError());
}

template <int PREC, int TP> void launch_tile(const AoMoLaunch& a, hipStream_t stream) {
    if (a.order == 0) launch_one<PREC, 0, kPanelsValue, TP>(a, stream);
    else if (a.order == 1) launch_one<PREC, 1, kPanelsGrad, TP>(a, stream);
    else if (a.hessian == 2) launch_one<PREC, 2, kPanelsAll, TP>(a, stream);
    else if (a.hessian == 3) launch_one<PREC, 2, kPanelsLap, TP>(a, stream);
    else launch_one<PREC, 2, kPanelsDiag, TP>(a, stream);
}

template <int PREC> void launch_prec(const AoMoLaunch& a, hipStream_t stream) {
    if constexpr (PREC == KUMO_HIP_FP64) {
        launch_tile<PREC, 16>(a, stream);
    } else {
        if (a.tile_points == 16) launch_tile<PREC, 16>(a, stream);
        else launch_tile<PREC, 32>(a, stream);
    }
}

// Folded rows of the fold workspace (one workgroup per screening block):
// row w = b * fold_per_block + j holds, at every MO slot, the sum over the
// kept pure AOs r of the entry of T[r][k] C[r] (pure-to-Cartesian transform
// times the MO coefficients, in r order as folded_row sums them), and the
// upper bounds of its norms.
template <int PREC>
__global__ void __launch_bounds__(256) fold_kernel(MolView m, MoOperand mo, KeptTables t) {
    const int b = blockIdx.x;
    const int n = t.fold_n[b];
    const int32_t* ent = t.ent + static_cast<long long>(b) * m.nshells;
    const uint32_t* ebits = t.ent_bits + static_cast<long long>(b) * m.nshells;
    for (int j = 0; j < n; ++j) {
        const int w = b * t.fold_per_block + j;
        const int code = t.fold_desc[w];
        const int c = kFoldedRow - code;
        const int q = c >> 5, k = c & 31;
        const DevShell& sh = m.shells[ent[q] & 0xffffff];
        const uint32_t bits = ebits[q];
        const double* tg = m.tg + sh.l * kMaxPure * kMaxCart + k;
        double* out = t.fold_rows + static_cast<long long>(w) * t.fold_slots;
        for (int slot = threadIdx.x; slot < mo.slots; slot += 256) {
            double acc = 0.0;
            for (int r = 0; r < sh.count; ++r) {
                const double tr = tg[r * kMaxCart];
                if (!((bits >> r) & 1u) || tr == 0.0) continue;
                acc += tr * mo.op64[static_cast<long long>(sh.prow + r) * mo.slots + slot];
            }
            out[slot] = acc;
            const long long at = static_cast<long long>(w) * t.fold_slots + slot;
            if constexpr (PREC == KUMO_HIP_FP32) {
                static_cast<float*>(t.fold_lo)[at] = static_cast<float>(acc);
            } else if constexpr (is_bf16<PREC>) {
                const long long plane = static_cast<long long>(t.fold_cap) * t.fold_slots;
#pragma unroll
                for (int sp = 0; sp < kBf16Planes; ++sp)
                    static_cast<uint16_t*>(t.fold_lo)[sp * plane + at] = bf16_term(acc, sp);
            }
        }
        if (threadIdx.x == 0) t.fold_norm[w] = folded_norm(m, mo, ent, ebits, code);
    }
}

} // namespace

void launch_fold(const MolView& m, const MoOperand& mo, const KeptTables& t, int nblk, int precision,
                 hipStream_t stream) {
    if (nblk <= 0 || t.fold_per_block <= 0) return;
    if (t.fold_slots != mo.slots || nblk * t.fold_per_block > t.fold_cap)
        invalid("fold workspace does not match the launch");
    switch (precision) {
    case KUMO_HIP_FP64: fold_kernel<KUMO_HIP_FP64><<<nblk, 256, 0, stream>>>(m, mo, t); break;
    case KUMO_HIP_FP32: fold_kernel<KUMO_HIP_FP32><<<nblk, 256, 0, stream>>>(m, mo, t); break;
    default: fold_kernel<KUMO_HIP_BF16X6><<<nblk, 256, 0, stream>>>(m, mo, t); break;
    }
    KUMO_HIP(hipGetLastError());
}

void launch_ao_mo(const AoMoLaunch& a, hipStream_t stream) {
    if (a.nblk <= 0 || a.slot1 <= a.slot0) return;
    if (a.slot0 % 16 || a.slot1 % 16) invalid("MO slot ranges must be multiples of 16");
    static_assert(Prec<KUMO_HIP_FP64>::T == kRescueTile, "rescue tiles are FP64 workgroup tiles");
    if (a.r_tile_blk && a.precision != KUMO_HIP_FP64) invalid("rescue tiles are evaluated in FP64");
    if (a.weights) {
        if (a.precision != KUMO_HIP_FP64 || a.order != 0 || a.r_tile_blk)
            invalid("weighted MO sums are FP64 values without rescue tiles");
        launch_one<KUMO_HIP_FP64, 0, kPanelsValue, 16, true>(a, stream);
        return;
    }
    if (a.tile_points == 64 && ao_mo_wave_supported(a)) {
        launch_ao_mo_wave(a, stream);
        return;
    }
    // (FP64 launches, the rescue's among them, take 16-point tiles whatever was asked)
    if (a.tile_points != 0 && a.tile_points != 16 && a.tile_points != 32 &&
        !(a.tile_points == 64 && a.precision == KUMO_HIP_FP64))
        invalid("tile points must be 16 or 32, or 64 for the split-bf16 precisions without full Hessians");
    switch (a.precision) {
    case KUMO_HIP_FP64: launch_prec<KUMO_HIP_FP64>(a, stream); break;
    case KUMO_HIP_FP32: launch_prec<KUMO_HIP_FP32>(a, stream); break;
    case KUMO_HIP_BF16X3: launch_prec<KUMO_HIP_BF16X3>(a, stream); break;
    default: launch_prec<KUMO_HIP_BF16X6>(a, stream); break;
    }
}

} // namespace kumo
