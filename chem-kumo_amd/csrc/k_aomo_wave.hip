// AO->MO contraction with wave-uniform AO generation, for the split-bf16
// precisions (density_derivatives_mo, src/density.cpp:234-351).
//
// One workgroup (8 waves, two per SIMD) owns 64 points of one screening block, one per
// lane, and up to 128 MO slots. Per contraction chunk (at most 32 operand
// rows, the mask kernel's chunk tables; a partially kept pure block enters as
// its Cartesian components with the kept pure AOs' coefficients folded when
// the chunk is staged):
//   1. the chunk's coefficient rows are staged into LDS as bf16 row pairs
//      (during the previous chunk's MFMA phase, into the other of two
//      buffers);
//   2. the chunk's entries (halves of the entries of six or more rows) are
//      taken by the waves in turn from an LDS counter, largest first; a wave
//      evaluates its rows at the 64 points of the tile, so the block's record,
//      angular momentum, primitive loop and component code are the same in
//      every lane (no divergence but the shell cutoff): the radial sums once
//      per (block, wave, point), then each Cartesian component with its
//      derivative panels, split into bf16 terms and written to LDS;
//   3. each wave multiplies one 16-point column of the tile by a quarter of
//      the MO tiles on MFMA (v_mfma_f32_16x16x32_bf16), accumulating in FP32.
// The chunk's plan (its entries' block records, row costs and the norms of
// its coefficient rows) is copied to LDS during the previous chunk's MFMA
// phase, so the generation reads it with LDS latency.
// The epilogue is the tile kernel's: FP64 dot products over the MOs of each
// tile, reduced over the tiles of each spin channel in a fixed order, and the
// error scales of the rescue summed over the waves in a fixed order.
//
// LDS: Phi as [plane][panel][row pair][point] 32-bit words (row 2q low,
// 2q + 1 high) with a line pitch of 68 words, so that the 64 lanes writing
// one row hit distinct banks and a B operand (rows 8g..8g+7 of one point:
// four words at a stride of 68) is read without bank conflicts; the
// coefficients as [plane][row pair][slot] words with a pitch of 132.
#include "kernels.hpp"

#include "../../include/kumo_hip.h"
#include "aomo_common.hpp"
#include "common.hpp"

#include <algorithm>
#include <cstddef>

namespace kumo {
namespace {
using namespace aomo;

constexpr int WT = 64;          // points per workgroup, one per lane
constexpr int WW = 8;           // waves per workgroup (two per SIMD)
constexpr int WTHR = WW * 64;   // threads per workgroup
constexpr int KC = kChunkRows;  // operand rows per chunk
constexpr int KQ = KC / 2;      // row pairs per chunk
constexpr int PP = WT + 4;      // words per Phi line
constexpr int CP = 132;         // words per coefficient line
constexpr int MTMAX = 8;        // MO tiles of a 128-slot group
constexpr int MW = WW / 4;      // waves sharing a 16-point column (MO tiles grp, grp + MW, ...)
constexpr int TPW = (MTMAX + MW - 1) / MW; // MO tiles per wave (at most)
constexpr int GMAX = 16;        // 16-byte granules of a 128-slot bf16 row
constexpr int RD = 30;          // doubles of a block record (ShellRec) in the plan
static_assert(sizeof(ShellRec) == RD * sizeof(double) && kRecPrim == 8, "plan record layout");
// Positions in a record copied to the plan (doubles, or ints for l, nprim, prim).
constexpr int kRecLnp = offsetof(ShellRec, l) / 8;      // int2 (l, nprim)
constexpr int kRecPrimInt = offsetof(ShellRec, prim) / 4;
constexpr int kRecE = offsetof(ShellRec, e) / 8;
constexpr int kRecW = offsetof(ShellRec, w) / 8;
constexpr int kRecSkip = offsetof(ShellRec, skip) / 8;
static_assert(offsetof(ShellRec, nprim) == offsetof(ShellRec, l) + 4 && offsetof(ShellRec, l) % 8 == 0 &&
                  offsetof(ShellRec, cutoff_r2) == 24 && offsetof(ShellRec, o) == 0,
              "record fields");

// Per-chunk data in LDS: the records of the current chunk's entries and the
// norms of its coefficient rows; and a ring of two chunks' indices (entry
// words, first rows, operand rows of the chunk rows) for the current and the
// next chunk, so that every global load of the next chunk's data has its
// address in LDS and can be issued before the current chunk's MFMA phase.
constexpr int PLAN_WORDS = KC * RD * 2 + KC * 2 + 2 * 3 * KC + 4;

template <int S, int NP> struct Layout {
    static constexpr int PHI = S * NP * KQ * PP;
    static constexpr int CS = S * KQ * CP;
    static constexpr int PLAN = PHI + 2 * CS; // word offset of the plan (after two coefficient buffers)
    // epilogue: red[MTMAX][kMoments][WT] doubles, then esl[WW][4][WT] floats
    static constexpr int RED = MTMAX * kMoments * WT * 2 + WW * 4 * WT;
    static constexpr int WORDS = PLAN + PLAN_WORDS > RED ? PLAN + PLAN_WORDS : RED;
};

// Writer of one lane's Phi values: each row's S bf16 terms of every panel go
// to the low (even row) or high (odd row) half of its row-pair word, as
// 16-bit stores, so rows need no pairing state and an entry may start or end
// at any row.
template <int S, int NP> struct PhiWriter {
    uint32_t* phi;
    int j;
    __device__ __attribute__((always_inline)) uint32_t* at(int sp, int p, int q) const {
        return phi + ((sp * NP + p) * KQ + q) * PP + j;
    }
    __device__ __attribute__((always_inline)) void put(int r, const double (&v)[NP]) const {
        uint16_t* base = reinterpret_cast<uint16_t*>(phi + (r >> 1) * PP + j) + (r & 1);
#pragma unroll
        for (int p = 0; p < NP; p += 2) {
            uint32_t t[S];
            split_pair<S>(v[p], p + 1 < NP ? v[p + 1] : 0.0, t);
#pragma unroll
            for (int sp = 0; sp < S; ++sp) {
                base[2 * ((sp * NP + p) * KQ * PP)] = static_cast<uint16_t>(t[sp] & 0xffffu);
                if (p + 1 < NP) base[2 * ((sp * NP + p + 1) * KQ * PP)] = static_cast<uint16_t>(t[sp] >> 16);
            }
        }
    }
};

__device__ __attribute__((always_inline)) inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// LDS plan of the kernel (word offsets from the plan base).
struct Plan {
    double* rec;   // [KC][RD] records of the current chunk's entries
    float2* nrm;   // [KC] norms of the current chunk's coefficient rows
    int* ring;     // [2][3][KC]: entry words, first rows, operand rows, by chunk parity
    int* next;     // work counter of the current chunk's generation
    __device__ __attribute__((always_inline)) int* went(int c) const { return ring + (c & 1) * 3 * KC; }
    __device__ __attribute__((always_inline)) int* wrow(int c) const { return ring + (c & 1) * 3 * KC + KC; }
    __device__ __attribute__((always_inline)) int* wop(int c) const { return ring + (c & 1) * 3 * KC + 2 * KC; }
};

// Global data of one chunk loaded into registers (one thread's share), so that
// the loads are in flight during the MFMA phase and written to LDS after it:
// a 16-byte piece of an entry record, two coefficient staging items (rows 2q
// and 2q + 1 of a granule), the norm of one coefficient row; and the ring
// indices of the chunk after it.
constexpr int kStageItems = (3 * KQ * GMAX + WTHR - 1) / WTHR; // coefficient staging items per thread
struct NextChunk {
    uint4 rec;
    uint4 cv[kStageItems][2];
    float2 nrm;
    int ient, irow, iop;
};

constexpr int REC_PIECES = RD * 8 / 16; // 16-byte pieces of a record

// Loads chunk cn's records, coefficient granules and norms (addresses from
// the ring) and chunk cn + 1's ring indices (when it exists).
template <int S>
__device__ __attribute__((always_inline)) inline void load_next(const AoMoLaunch& a, const Plan& pl, int cn,
                                                              int nent, int nchunks, int e1, int e2,
                                                              const int32_t* ent, const uint8_t* erow,
                                                              const int32_t* crows, int slot_lo, int G, int tid,
                                                              NextChunk& nx) {
    // Every load is issued unconditionally from a valid address (indices
    // clamped into range; store_next drops what is not needed), so that no
    // branch separates the loads from the MFMA phase that hides them.
    const MolView& m = a.mol;
    const int* went = pl.went(cn);
    const int* wop = pl.wop(cn);
    {
        const int i = min(tid / REC_PIECES, nent - 1), f = tid % REC_PIECES;
        nx.rec = reinterpret_cast<const uint4*>(&m.rec[went[i] & 0xffffff])[f];
    }
    const uint16_t* op = a.mo.opbf;
    const long long op_plane = static_cast<long long>(a.mo.rows) * a.mo.slots;
#pragma unroll
    for (int u = 0; u < kStageItems; ++u) {
        const int item = min(tid + WTHR * u, S * KQ * GMAX - 1);
        const int gi = min(item % GMAX, G - 1);
        const int q = (item / GMAX) % KQ;
        const int sp = item / (KQ * GMAX);
        const uint16_t* src = op + sp * op_plane + slot_lo + gi * 8;
        const uint16_t* ws = static_cast<const uint16_t*>(a.tables.fold_lo) +
                             static_cast<long long>(sp) * a.tables.fold_cap * a.tables.fold_slots + slot_lo + gi * 8;
        // operand table rows, or fold workspace rows as staged
        const auto row = [&](int g) __attribute__((always_inline)) {
            return g <= kFoldedWs ? ws + static_cast<long long>(kFoldedWs - g) * a.tables.fold_slots
                                  : src + static_cast<long long>(max(g, 0)) * a.mo.slots;
        };
        nx.cv[u][0] = *reinterpret_cast<const uint4*>(row(wop[2 * q]));
        nx.cv[u][1] = *reinterpret_cast<const uint4*>(row(wop[2 * q + 1]));
    }
    nx.nrm = a.mo.norm[max(wop[tid & (KC - 1)], 0)];
    // ring indices of chunk cn + 1: lanes of wave 0 (entries), wave 1 (first
    // rows), wave 2 (operand rows); e1 < e2 when the chunk exists
    if (cn + 1 < nchunks) {
        const int l = tid & 63;
        const int ei = e1 + min(l, e2 - e1 - 1);
        nx.ient = ent[ei];
        nx.irow = erow[ei];
        nx.iop = crows[2 * ((cn + 1) * KC + (l & (KC - 1)))];
    }
}

// Writes what load_next loaded: chunk cn's records, coefficients (into cs)
// and norms, taking folded rows (partially kept pure blocks, rare) from the
// fold workspace or forming them here, and chunk cn + 1's ring indices.
template <int S>
__device__ __attribute__((always_inline)) inline void store_next(const AoMoLaunch& a, const Plan& pl, int cn,
                                                               int nent, int nchunks, int e1, int e2,
                                                               const int32_t* ent, const uint32_t* ebits,
                                                               int slot_lo, int G, int tid, uint32_t* cs,
                                                               const NextChunk& nx) {
    const int* wop = pl.wop(cn);
    if (tid == 0) *pl.next = 0; // the generation of chunk cn starts from its first work unit
    if (tid < nent * REC_PIECES) reinterpret_cast<uint4*>(pl.rec)[tid] = nx.rec;
#pragma unroll
    for (int u = 0; u < kStageItems; ++u) {
        const int item = tid + WTHR * u;
        if (item >= S * KQ * GMAX) continue;
        const int gi = item % GMAX;
        const int q = (item / GMAX) % KQ;
        const int sp = item / (KQ * GMAX);
        if (gi >= G) continue;
        const int g0 = wop[2 * q], g1 = wop[2 * q + 1];
        const auto loaded = [](int g) { return g >= 0 || g <= kFoldedWs; };
        uint4 v0 = loaded(g0) ? nx.cv[u][0] : uint4{0, 0, 0, 0}, v1 = loaded(g1) ? nx.cv[u][1] : uint4{0, 0, 0, 0};
        if (g0 < -1 && g0 > kFoldedWs) v0 = folded_bf16(a.mol, a.mo, ent, ebits, g0, slot_lo + gi * 8, sp);
        if (g1 < -1 && g1 > kFoldedWs) v1 = folded_bf16(a.mol, a.mo, ent, ebits, g1, slot_lo + gi * 8, sp);
        const auto lo = [](uint32_t x, uint32_t y) { return (x & 0xffffu) | (y << 16); };
        const auto hi = [](uint32_t x, uint32_t y) { return (x >> 16) | (y & 0xffff0000u); };
        uint4* d = reinterpret_cast<uint4*>(cs + (sp * KQ + q) * CP + gi * 8);
        d[0] = uint4{lo(v0.x, v1.x), hi(v0.x, v1.x), lo(v0.y, v1.y), hi(v0.y, v1.y)};
        d[1] = uint4{lo(v0.z, v1.z), hi(v0.z, v1.z), lo(v0.w, v1.w), hi(v0.w, v1.w)};
    }
    if (tid >= WTHR - KC) {
        const int r = tid - (WTHR - KC);
        const int g = wop[r];
        pl.nrm[r] = g <= kFoldedWs ? a.tables.fold_norm[kFoldedWs - g]
                    : (g < -1 ? folded_norm(a.mol, a.mo, ent, ebits, g) : (g >= 0 ? nx.nrm : float2{0.f, 0.f}));
    }
    if (cn + 1 < nchunks && tid < 3 * 64) {
        const int l = tid & 63, w = tid >> 6;
        if (w == 0 && l < e2 - e1) pl.went(cn + 1)[l] = nx.ient;
        if (w == 1 && l < e2 - e1) pl.wrow(cn + 1)[l] = nx.irow;
        if (w == 2 && l < KC) pl.wop(cn + 1)[l] = nx.iop;
    }
}

// Radial sums R0, R1, R2 of a plan entry at r2 (ao_eval.hpp radial_sums_rec),
// leaving out primitives beyond their skip radius. The next primitive's data
// are read while the current exponential is evaluated.
template <int ORDER>
__device__ __attribute__((always_inline)) inline void radial_sums_plan(const double* pd, int np, int prim,
                                                                     const MolView& m, double r2, bool inside,
                                                                     double (&R)[3]) {
#pragma clang fp contract(fast)
    double r0 = 0.0, r1 = 0.0, rr = 0.0;
    const auto add = [&](double a, double u) __attribute__((always_inline)) {
        const double g = u * exp(-a * r2);
        r0 += g;
        if constexpr (ORDER >= 1) {
            const double t = -2.0 * a * g;
            r1 += t;
            if constexpr (ORDER == 2) rr += -2.0 * a * t;
        }
    };
    if (np <= kRecPrim) {
        double an = pd[kRecE], wn = pd[kRecW], sn = pd[kRecSkip];
#pragma unroll 1
        for (int p = 0; p < np; ++p) {
            const double ap = an, wp = wn, sp = sn;
            if (p + 1 < np) {
                an = pd[kRecE + 1 + p];
                wn = pd[kRecW + 1 + p];
                sn = pd[kRecSkip + 1 + p];
            }
            if (inside && r2 <= sp) add(ap, wp);
        }
    } else if (inside) {
#pragma unroll 1
        for (int p = 0; p < np; ++p)
            add(m.exps[prim + p], m.radw[prim + p]);
    }
    R[0] = r0;
    R[1] = r1;
    R[2] = rr;
}

// Components [ra, rb) of plan entry i at this lane's point, as chunk rows
// r0 + ra .. r0 + rb - 1.
template <int S, int ORDER, unsigned PSET>
__device__ __attribute__((always_inline)) inline void gen_entry(const MolView& m, const double* pd, int r0,
                                                              int l, int np, int ra, int rb, const float2* pn,
                                                              double px, double py, double pz, bool valid,
                                                              const PhiWriter<S, popcount10(PSET)>& w,
                                                              float (&es)[4], bool scale) {
    constexpr int NP = popcount10(PSET);
    constexpr bool kGradScale = ORDER >= 1 && (PSET & 0xfu) == 0xfu;
    const double x = px - pd[0], y = py - pd[1], z = pz - pd[2];
    const double r2 = sq_norm(x, y, z);
    const bool inside = valid && r2 <= pd[3];
    double R[3];
    radial_sums_plan<ORDER>(pd, np, reinterpret_cast<const int*>(pd)[kRecPrimInt], m, r2, inside, R);
    const auto emit = [&](int k, const double (&v)[NP]) __attribute__((always_inline)) {
        if (k < ra || k >= rb) return; // another wave's rows
        const int r = r0 + k;
        if (scale) {
            // error scales from the value and gradient panels
            const float2 nn = pn[r];
            const float av = static_cast<float>(v[0]);
            const float wx2 = nn.x * nn.x, wy2 = nn.y * nn.y;
            es[0] += (av * av) * wx2;
            es[1] += (av * av) * wy2;
            if constexpr (kGradScale) {
                const float gx = static_cast<float>(v[1]), gy = static_cast<float>(v[2]),
                            gz = static_cast<float>(v[3]);
                const float ag2 = gx * gx + gy * gy + gz * gz;
                es[2] += ag2 * wx2;
                es[3] += ag2 * wy2;
            }
        }
        w.put(r, v);
    };
    switch (l) {
    case 0: cart_components<0, ORDER, PSET>(x, y, z, R, emit); break;
    case 1: cart_components<1, ORDER, PSET>(x, y, z, R, emit); break;
    case 2: cart_components<2, ORDER, PSET>(x, y, z, R, emit); break;
    case 3: cart_components<3, ORDER, PSET>(x, y, z, R, emit); break;
    default: {
        // g and h blocks: components one at a time (runtime powers)
        const int nc = (l + 1) * (l + 2) / 2;
        for (int c = 0; c < nc; ++c) {
            double v[NP];
            cart_component_rt<ORDER, PSET>(l, c, x, y, z, R, v);
            emit(c, v);
        }
    }
    }
}

// One 64-point tile tx of the launch (workgroup-uniform).
template <int PREC, int ORDER, unsigned PSET>
__device__ __attribute__((always_inline)) inline void wave_tile(const AoMoLaunch& a, const int tx) {
    constexpr int S = PREC == KUMO_HIP_BF16X6 ? 3 : 2;
    constexpr int NP = popcount10(PSET);
    using Lay = Layout<S, NP>;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    uint32_t* phi = reinterpret_cast<uint32_t*>(smem);
    uint32_t* cs2 = phi + Lay::PHI; // two coefficient buffers, chunks alternating
    const Plan pl{reinterpret_cast<double*>(phi + Lay::PLAN), reinterpret_cast<float2*>(phi + Lay::PLAN + KC * RD * 2),
                  reinterpret_cast<int*>(phi + Lay::PLAN + KC * RD * 2 + KC * 2),
                  reinterpret_cast<int*>(phi + Lay::PLAN + KC * RD * 2 + KC * 2 + 2 * 3 * KC)};
    double* red = reinterpret_cast<double*>(smem);
    float* esl = reinterpret_cast<float*>(smem) + MTMAX * kMoments * WT * 2;

    const int tid = threadIdx.x, lane = tid & 63, wave = uniform(tid >> 6);
    const int tiles_per_block = (a.block_size + WT - 1) / WT;
    const int blk_local = tx / tiles_per_block;
    const int p0 = (tx % tiles_per_block) * WT;
    const long long blk = a.first_block + blk_local;
    const long long bstart = blk * a.block_size;
    const int bcount = static_cast<int>(min(static_cast<long long>(a.block_size), a.npoints - bstart));
    if (p0 >= bcount) return; // uniform over the workgroup
    const int nt_valid = min(WT, bcount - p0);
    const int slot_lo = a.slot0 + 128 * blockIdx.y;
    const int width = min(a.slot1, slot_lo + 128) - slot_lo;
    const int MT = width / 16;
    const bool scale = a.rho_scale != nullptr;

    const MolView& m = a.mol;
    const int nsh = m.nshells;
    const int32_t* ent = a.tables.ent + static_cast<long long>(blk_local) * nsh;
    const uint32_t* ebits = a.tables.ent_bits + static_cast<long long>(blk_local) * nsh;
    const uint8_t* erow = a.tables.ent_row + static_cast<long long>(blk_local) * nsh;
    const int32_t* cb = a.tables.chunk_begin + static_cast<long long>(blk_local) * (a.tables.max_chunks + 1);
    const int32_t* crows = a.tables.chunk_rows + static_cast<long long>(blk_local) * a.tables.max_chunks * KC * 2;
    const int nchunks = a.tables.nchunks[blk_local];

    const int j = lane;
    const bool valid = j < nt_valid;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (valid) point_coords_from(a.ps, bstart + p0, j, px, py, pz);

    f32x4 acc[TPW][NP];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int i = 0; i < NP; ++i)
            acc[t][i] = f32x4{0, 0, 0, 0};
    float es[4] = {0.f, 0.f, 0.f, 0.f};

    const int G = width / 8; // granules per coefficient row
    const int g = lane >> 4, r16 = lane & 15;
    const int nt = wave & 3, grp = wave >> 2;
    const PhiWriter<S, NP> wr{phi, j};

    // chunk starts: lane i holds cb[i] (the chunks past lane 63 read it from memory)
    const int cbv = cb[min(lane, nchunks)];
    const auto cbeg = [&](int c) __attribute__((always_inline)) {
        return c < 64 ? __builtin_amdgcn_readlane(cbv, c) : uniform(cb[c]);
    };
    if (nchunks > 0) {
        // ring indices of chunk 0 (loads from clamped, valid indices), then
        // its records, coefficients and norms
        {
            const int e0 = cbeg(0), e1 = cbeg(1);
            const int l = tid & 63, w = tid >> 6;
            const int ei = e0 + min(l, e1 - e0 - 1);
            const int ve = ent[ei], vr = erow[ei], vo = crows[2 * (l & (KC - 1))];
            if (w == 0 && l < e1 - e0) pl.went(0)[l] = ve;
            if (w == 1 && l < e1 - e0) pl.wrow(0)[l] = vr;
            if (w == 2 && l < KC) pl.wop(0)[l] = vo;
        }
        __syncthreads();
        NextChunk nx;
        const int e0 = cbeg(0), e1 = cbeg(1), e2 = nchunks > 1 ? cbeg(2) : e1;
        load_next<S>(a, pl, 0, e1 - e0, nchunks, e1, e2, ent, erow, crows, slot_lo, G, tid, nx);
        store_next<S>(a, pl, 0, e1 - e0, nchunks, e1, e2, ent, ebits, slot_lo, G, tid, cs2, nx);
        __syncthreads();
    }
    for (int c = 0; c < nchunks; ++c) {
        const int eb0 = cbeg(c), eb1 = cbeg(c + 1);
        const int nent = eb1 - eb0;
        // 2. Phi: the chunk's work units (an entry, or half of an entry of six
        //    or more rows) taken by the waves in turn from an LDS counter, in
        //    the chunk's descending-l order (largest first), so that no wave
        //    waits for a static share that turned out longer than the others'.
        {
            int l_i = 0, np_i = 0, r0_i = 0, nu_i = 0;
            if (lane < nent) {
                const int2 lnp = reinterpret_cast<const int2*>(pl.rec + lane * RD)[kRecLnp];
                l_i = lnp.x;
                np_i = lnp.y;
                r0_i = pl.wrow(c)[lane];
                nu_i = (l_i + 1) * (l_i + 2) / 2 >= 6 ? 2 : 1;
            }
            int incl = nu_i;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o, 64);
                if (lane >= o) incl += t;
            }
            const int nunits = __builtin_amdgcn_readlane(incl, 63);
            const int first_u = incl - nu_i; // first unit of entry `lane`
            for (;;) {
                int u = 0;
                if (lane == 0) u = __hip_atomic_fetch_add(pl.next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                u = uniform(u);
                if (u >= nunits) break;
                // the entry of unit u: the last lane whose first unit is <= u
                const unsigned long long le = __ballot(lane < nent && first_u <= u);
                const int e = 63 - __builtin_clzll(le);
                const int l = __builtin_amdgcn_readlane(l_i, e);
                const int np = __builtin_amdgcn_readlane(np_i, e);
                const int r0 = __builtin_amdgcn_readlane(r0_i, e);
                const int n = (l + 1) * (l + 2) / 2;
                const int part = u - __builtin_amdgcn_readlane(first_u, e);
                const int nu = __builtin_amdgcn_readlane(nu_i, e);
                const int ra = nu == 2 && part == 1 ? n / 2 : 0;
                const int rb = nu == 2 && part == 0 ? n / 2 : n;
                gen_entry<S, ORDER, PSET>(m, pl.rec + e * RD, r0, l, np, ra, rb, pl.nrm, px, py, pz, valid, wr, es,
                                          scale);
            }
            // rows past the chunk's end are zero
            if (wave == WW - 1) {
                const int last_r0 = __builtin_amdgcn_readlane(r0_i, nent - 1);
                const int last_l = __builtin_amdgcn_readlane(l_i, nent - 1);
                const int used = last_r0 + (last_l + 1) * (last_l + 2) / 2;
                for (int q = used >> 1; q < KQ; ++q)
#pragma unroll
                    for (int p = 0; p < NP; ++p)
#pragma unroll
                        for (int sp = 0; sp < S; ++sp) {
                            uint32_t* d = wr.at(sp, p, q);
                            if (q == (used >> 1) && (used & 1)) reinterpret_cast<uint16_t*>(d)[1] = 0;
                            else *d = 0u;
                        }
            }
        }
        __syncthreads();
        // 3. MFMA: this wave's 16-point column nt and MO tiles half + 2t. The
        //    next chunk's data are loaded before and written to LDS after it
        //    (the generation of this chunk, the only reader of the records,
        //    norms and ring slot they replace, has passed the barrier).
        NextChunk nx;
        const bool more = c + 1 < nchunks;
        const int e2 = more ? cbeg(c + 2) : eb1;
        const int e3 = c + 2 < nchunks ? cbeg(c + 3) : e2;
        if (more) load_next<S>(a, pl, c + 1, e2 - eb1, nchunks, e2, e3, ent, erow, crows, slot_lo, G, tid, nx);
        {
            const uint32_t* cs = cs2 + (c & 1) * Lay::CS;
            const int jb = nt * 16 + r16;
            // two MO tiles at a time (their A operands held, the B operands of
            // each panel read once per pair, one panel ahead of its MFMAs)
            const auto pair = [&]<bool TWO>(int t0) __attribute__((always_inline)) {
                bf16x8 av[2][S];
#pragma unroll
                for (int t = 0; t < (TWO ? 2 : 1); ++t) {
                    const int mt = grp + MW * (t0 + t);
#pragma unroll
                    for (int sp = 0; sp < S; ++sp) {
                        const uint32_t* w = cs + (sp * KQ + 4 * g) * CP + mt * 16 + r16;
                        av[t][sp] = __builtin_bit_cast(bf16x8, uint4{w[0], w[CP], w[2 * CP], w[3 * CP]});
                    }
                }
                const auto load_b = [&](int p, bf16x8 (&bv)[S]) __attribute__((always_inline)) {
#pragma unroll
                    for (int sp = 0; sp < S; ++sp) {
                        const uint32_t* w = phi + ((sp * NP + p) * KQ + 4 * g) * PP + jb;
                        bv[sp] = __builtin_bit_cast(bf16x8, uint4{w[0], w[PP], w[2 * PP], w[3 * PP]});
                    }
                };
                bf16x8 bv[2][S];
                load_b(0, bv[0]);
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    if (p + 1 < NP) load_b(p + 1, bv[(p + 1) & 1]);
                    const bf16x8(&b)[S] = bv[p & 1];
                    // small terms first; the tiles' chains interleaved
                    const auto mfma = [&](int ia, int ib) __attribute__((always_inline)) {
                        acc[t0][p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][ia], b[ib], acc[t0][p], 0, 0, 0);
                        if constexpr (TWO)
                            acc[t0 + 1][p] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1][ia], b[ib], acc[t0 + 1][p], 0, 0, 0);
                    };
                    if constexpr (S == 3) {
                        mfma(2, 0);
                        mfma(1, 1);
                        mfma(0, 2);
                    }
                    mfma(1, 0);
                    mfma(0, 1);
                    mfma(0, 0);
                }
            };
            const int nmine = MT > grp ? (MT - grp + MW - 1) / MW : 0; // MO tiles grp, grp + MW, ... below MT
#pragma unroll
            for (int t0 = 0; t0 < TPW; t0 += 2) {
                if constexpr (true) {
                    if (t0 + 1 < TPW && nmine >= t0 + 2) pair.template operator()<true>(t0);
                    else if (nmine >= t0 + 1) pair.template operator()<false>(t0);
                }
            }
        }
        if (more)
            store_next<S>(a, pl, c + 1, e2 - eb1, nchunks, e2, e3, ent, ebits, slot_lo, G, tid,
                          cs2 + ((c + 1) & 1) * Lay::CS, nx);
        __syncthreads();
    }

    // Epilogue (red and esl alias the Phi and coefficient buffers).
    if (scale) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            esl[(wave * 4 + q) * WT + j] = es[q];
    }
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int mt = grp + MW * t;
        if (mt >= MT) break;
        double q[kMoments];
#pragma unroll
        for (int k = 0; k < kMoments; ++k)
            q[k] = 0.0;
        const auto V = [&](int comp, int r) __attribute__((always_inline)) -> double {
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
                    red[(static_cast<size_t>(mt) * kMoments + k) * WT + nt * 16 + r16] = q[k];
    }
    __syncthreads();
    const int alpha_tiles_end = max(0, min(MT, (a.mo.alpha_slots - slot_lo) / 16));
    for (int item = tid; item < 2 * kMoments * WT; item += WTHR) {
        const int jj = item % WT;
        const int k = (item / WT) % kMoments;
        const int ch = item / (WT * kMoments);
        if (jj >= nt_valid) continue;
        const int mt0 = ch == 0 ? 0 : alpha_tiles_end, mt1 = ch == 0 ? alpha_tiles_end : MT;
        double s = 0.0;
        if (moment_used<ORDER, PSET>(k))
            for (int mt = mt0; mt < mt1; ++mt)
                s += red[(static_cast<size_t>(mt) * kMoments + k) * WT + jj];
        const double f = k == 0 ? 1.0 : (k <= 3 ? 2.0 : (k <= 6 ? 0.5 : 2.0));
        const long long point = bstart + p0 + jj - a.batch_first;
        double* dst = a.moments + (static_cast<long long>(blockIdx.y) * 2 + ch) * kMoments * a.batch_points;
        dst[k * a.batch_points + point] = f * s;
        if (k == 0 && scale && blockIdx.y == 0) {
            float e = 0.f, eg = 0.f;
#pragma unroll
            for (int w = 0; w < WW; ++w) {
                e += esl[(w * 4 + ch) * WT + jj];
                eg += esl[(w * 4 + 2 + ch) * WT + jj];
            }
            a.rho_scale[ch * a.batch_points + point] = sqrtf(e);
            if (ORDER >= 1 && (PSET & 0xfu) == 0xfu) a.rho_scale[(2 + ch) * a.batch_points + point] = sqrtf(eg);
        }
    }
}

// Persistent launch: each workgroup (one per CU: its LDS and registers leave
// room for no second) takes the tiles in turn, so the per-workgroup start-up
// is paid once per CU rather than once per tile.
template <int PREC, int ORDER, unsigned PSET>
__global__ void __launch_bounds__(WTHR, WW / 4) aomo_wave_kernel(AoMoLaunch a, int ntiles) {
    for (int tx = blockIdx.x; tx < ntiles; tx += gridDim.x) {
        wave_tile<PREC, ORDER, PSET>(a, tx);
        __syncthreads(); // the next tile reuses the LDS
    }
}

// Compute units of the current device (cached per device).
int device_cus() {
    static int cus[64] = {};
    int dev = 0;
    KUMO_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    if (!cus[dev]) {
        int n = 0;
        KUMO_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        cus[dev] = n > 0 ? n : 1;
    }
    return cus[dev];
}

template <int PREC, int ORDER, unsigned PSET> void launch_one(const AoMoLaunch& a, hipStream_t stream) {
    constexpr int S = PREC == KUMO_HIP_BF16X6 ? 3 : 2;
    const size_t lds = sizeof(uint32_t) * Layout<S, popcount10(PSET)>::WORDS;
    static_assert(sizeof(uint32_t) * Layout<3, 7>::WORDS <= 160 * 1024, "LDS of the largest panel set");
    const int tiles_per_block = (a.block_size + WT - 1) / WT;
    const int groups = (a.slot1 - a.slot0 + 127) / 128;
    const int ntiles = a.nblk * tiles_per_block;
    const dim3 grid(static_cast<unsigned>(std::min(ntiles, device_cus())), static_cast<unsigned>(groups));
    static bool configured = false;
    if (!configured) {
        KUMO_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&aomo_wave_kernel<PREC, ORDER, PSET>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        configured = true;
    }
    aomo_wave_kernel<PREC, ORDER, PSET><<<grid, dim3(WTHR), lds, stream>>>(a, ntiles);
    KUMO_HIP(hipGetLastError());
}

template <int PREC> void launch_prec(const AoMoLaunch& a, hipStream_t stream) {
    if (a.order == 0) launch_one<PREC, 0, kPanelsValue>(a, stream);
    else if (a.order == 1) launch_one<PREC, 1, kPanelsGrad>(a, stream);
    else if (a.hessian == 3) launch_one<PREC, 2, kPanelsLap>(a, stream);
    else launch_one<PREC, 2, kPanelsDiag>(a, stream);
}

} // namespace

bool ao_mo_wave_supported(const AoMoLaunch& a) {
    return (a.precision == KUMO_HIP_BF16X6 || a.precision == KUMO_HIP_BF16X3) && a.hessian != 2 &&
           a.r_tile_blk == nullptr && a.weights == nullptr;
}

void launch_ao_mo_wave(const AoMoLaunch& a, hipStream_t stream) {
    if (!ao_mo_wave_supported(a)) invalid("the wave-uniform contraction takes split-bf16 precisions without full Hessians");
    if (a.precision == KUMO_HIP_BF16X6) launch_prec<KUMO_HIP_BF16X6>(a, stream);
    else launch_prec<KUMO_HIP_BF16X3>(a, stream);
}

} // namespace kumo
