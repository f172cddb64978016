// Screening masks (src/ao_screening.hpp:35-50 as applied per point block in
// src/density.cpp:454-467) and AO panel evaluation.
//
// One workgroup per screening block: each lane evaluates the AO blocks at one
// point, the largest |value| over the screened panels is reduced across the
// wave with shuffles and across the workgroup with a 64-bit LDS atomic max on
// the IEEE bit pattern (monotone for non-negative doubles). The kept AO list,
// the kept AO blocks and their packing into contraction chunks are written
// for the contraction kernels. Built with -ffp-contract=off (see ao_eval.hpp).
#include "kernels.hpp"

#include "common.hpp"

namespace kumo {
namespace {

constexpr int kMaskThreads = 128; // one lane per point of a default 128-point block

template <int ORDER>
__global__ void __launch_bounds__(kMaskThreads)
    mask_kernel(MolView m, PointSource ps, long long npoints, int block_size, long long first_block, double threshold,
                KeptTables t) {
    // LDS: smax[nao], then per AO block kept bits, entries, entry slots, and
    // the rows used by each chunk.
    extern __shared__ unsigned long long smax[];
    const int tid = threadIdx.x;
    const long long blk = first_block + blockIdx.x;
    const long long start = blk * block_size;
    const int count = static_cast<int>(min(static_cast<long long>(block_size), npoints - start));
    const int nao = m.nao;
    for (int a = tid; a < nao; a += kMaskThreads)
        smax[a] = 0ull;
    __syncthreads();
    constexpr unsigned PSET = ORDER == 0 ? kPanelsValue : (ORDER == 1 ? kPanelsGrad : kPanelsAll);
    constexpr int NP = popcount10(PSET);
    const bool screen = threshold != 0.0;
    if (screen) {
        for (int p0 = 0; p0 < count; p0 += kMaskThreads) {
            const int j = p0 + tid;
            const bool active = j < count;
            double px = 0, py = 0, pz = 0;
            if (active) point_coords(ps, start + j, px, py, pz);
            for (int s = 0; s < m.nshells; ++s) {
                const DevShell sh = m.shells[s];
                const double x = px - sh.o[0], y = py - sh.o[1], z = pz - sh.o[2];
                const double r2 = sq_norm(x, y, z);
                const bool inside = active && r2 <= sh.cutoff_r2;
                if (!__any(inside)) continue;
                double amax[kMaxCart];
#pragma unroll
                for (int r = 0; r < kMaxCart; ++r)
                    amax[r] = 0.0;
                if (inside) {
                    eval_block<ORDER, PSET>(sh, m, x, y, z, r2, [&](int row, const double (&v)[NP]) __attribute__((always_inline)) {
                        double a = 0.0;
#pragma unroll
                        for (int i = 0; i < NP; ++i)
                            a = fmax(a, fabs(v[i]));
                        amax[row] = a;
                    });
                }
#pragma unroll
                for (int r = 0; r < kMaxCart; ++r) {
                    if (r < sh.count) {
                        const double v = wave_max(amax[r]);
                        if ((threadIdx.x & 63) == 0 && v > 0.0)
                            atomicMax(&smax[sh.ao + r], static_cast<unsigned long long>(__double_as_longlong(v)));
                    }
                }
            }
        }
    }
    __syncthreads();
    // Flags and the ascending kept list (block-wide scan over the AOs).
    uint8_t* kept = t.kept + static_cast<long long>(blockIdx.x) * nao;
    int32_t* list = t.kept_list + static_cast<long long>(blockIdx.x) * nao;
    __shared__ int wave_counts[kMaskThreads / 64];
    __shared__ int base;
    if (tid == 0) base = 0;
    __syncthreads();
    for (int a0 = 0; a0 < nao; a0 += kMaskThreads) {
        const int a = a0 + tid;
        bool k = false;
        if (a < nao) {
            k = !screen || count == 0 || __longlong_as_double(static_cast<long long>(smax[a])) > threshold;
            kept[a] = k ? 1 : 0;
        }
        const unsigned long long ballot = __ballot(k);
        const int lane = tid & 63, wave = tid >> 6;
        const int before = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0) wave_counts[wave] = __popcll(ballot);
        __syncthreads();
        int offset = base;
        for (int w = 0; w < wave; ++w)
            offset += wave_counts[w];
        if (k) list[offset + before] = a;
        __syncthreads();
        if (tid == 0) {
            int total = 0;
            for (int w = 0; w < kMaskThreads / 64; ++w)
                total += wave_counts[w];
            base += total;
        }
        __syncthreads();
    }
    if (tid == 0) t.nkept[blockIdx.x] = base;
    // Kept bits of every AO block.
    const int nsh = m.nshells;
    uint32_t* sbits = reinterpret_cast<uint32_t*>(smax + nao);
    int32_t* sent = reinterpret_cast<int32_t*>(sbits + nsh);
    int32_t* sslot = sent + nsh;
    int32_t* sused = sslot + nsh;
    for (int s = tid; s < nsh; s += kMaskThreads) {
        const DevShell& sh = m.shells[s];
        uint32_t bits = 0;
        for (int r = 0; r < sh.count; ++r)
            if (!screen || count == 0 || __longlong_as_double(static_cast<long long>(smax[sh.ao + r])) > threshold)
                bits |= 1u << r;
        sbits[s] = bits;
    }
    __syncthreads();
    // Ordered compaction of the kept blocks: pass 0 in AO order (AO-P
    // tables), passes l = max..0 by descending l (AO-MO entries).
    int32_t* ks = t.kept_shell + static_cast<long long>(blockIdx.x) * nsh;
    uint32_t* kb = t.kept_bits + static_cast<long long>(blockIdx.x) * nsh;
    int nent = 0;
    for (int pass = -1; pass <= kMaxL; ++pass) {
        const int want_l = pass < 0 ? -1 : kMaxL - pass;
        if (tid == 0) base = 0;
        __syncthreads();
        for (int s0 = 0; s0 < nsh; s0 += kMaskThreads) {
            const int s = s0 + tid;
            bool take = false;
            uint32_t bits = 0;
            if (s < nsh) {
                bits = sbits[s];
                take = bits != 0 && (want_l < 0 || m.shells[s].l == want_l);
            }
            const unsigned long long ballot = __ballot(take);
            const int lane = tid & 63, wave = tid >> 6;
            const int before = __popcll(ballot & ((1ull << lane) - 1ull));
            if (lane == 0) wave_counts[wave] = __popcll(ballot);
            __syncthreads();
            int offset = base;
            for (int w = 0; w < wave; ++w)
                offset += wave_counts[w];
            if (take) {
                const int pos = offset + before;
                if (want_l < 0) {
                    ks[pos] = s;
                    kb[pos] = bits;
                } else {
                    const DevShell& sh = m.shells[s];
                    const bool full = bits == (sh.count >= 32 ? 0xffffffffu : ((1u << sh.count) - 1u));
                    const int mode = full ? kEntryCart : (sh.pure ? kEntryPure : kEntryCartMasked);
                    sent[nent + pos] = s | (mode << 24);
                    sslot[nent + pos] = mode == kEntryPure ? sh.count : sh.ncart; // rows, packed below
                }
            }
            __syncthreads();
            if (tid == 0) {
                int total = 0;
                for (int w = 0; w < kMaskThreads / 64; ++w)
                    total += wave_counts[w];
                base += total;
            }
            __syncthreads();
        }
        if (want_l < 0) {
            if (tid == 0) t.nkshell[blockIdx.x] = base;
        } else {
            nent += base;
        }
        __syncthreads();
    }
    // Next-fit packing of the entries into chunks of at most kChunkRows rows.
    int32_t* cb = t.chunk_begin + static_cast<long long>(blockIdx.x) * (t.max_chunks + 1);
    __shared__ int packed_entries;
    if (tid == 0) {
        int nc = 0, rows = kChunkRows, q = 0; // forces a new chunk at the first entry
        for (; q < nent; ++q) {
            const int n = sslot[q];
            if (rows + n > kChunkRows) {
                if (nc > 0) sused[nc - 1] = rows;
                if (nc == t.max_chunks) break; // not reached: max_chunks bounds every block
                cb[nc++] = q;
                rows = 0;
            }
            sslot[q] = ((nc - 1) << 8) | rows;
            rows += n;
        }
        if (nc > 0) sused[nc - 1] = rows;
        cb[nc] = q;
        t.nchunks[blockIdx.x] = nc;
        base = nc;
        packed_entries = q;
    }
    __syncthreads();
    const int nc = base;
    nent = packed_entries;
    int32_t* en = t.ent + static_cast<long long>(blockIdx.x) * nsh;
    uint32_t* eb = t.ent_bits + static_cast<long long>(blockIdx.x) * nsh;
    uint8_t* er = t.ent_row + static_cast<long long>(blockIdx.x) * nsh;
    int32_t* crows = t.chunk_rows + static_cast<long long>(blockIdx.x) * t.max_chunks * kChunkRows;
    for (int q = tid; q < nent; q += kMaskThreads) {
        const int e = sent[q];
        const int s = e & 0xffffff, mode = e >> 24;
        const int c = sslot[q] >> 8, row0 = sslot[q] & 0xff;
        const DevShell& sh = m.shells[s];
        en[q] = e;
        eb[q] = sbits[s];
        er[q] = static_cast<uint8_t>(row0);
        const int first = mode == kEntryPure ? sh.prow : sh.crow;
        const int n = mode == kEntryPure ? sh.count : sh.ncart;
        // Rows of dropped AOs (partially kept blocks) get zero coefficients.
        const uint32_t bits = sbits[s];
        if (c < t.max_chunks)
            for (int i = 0; i < n; ++i)
                crows[c * kChunkRows + row0 + i] = (mode == kEntryCart || ((bits >> i) & 1u)) ? first + i : -1;
    }
    for (int item = tid; item < nc * kChunkRows; item += kMaskThreads) {
        const int c = item / kChunkRows, r = item % kChunkRows;
        if (r >= sused[c]) crows[item] = -1;
    }
}

template <int ORDER>
__global__ void __launch_bounds__(128) ao_eval_kernel(MolView m, PointSource ps, long long n, double* out) {
    const long long j = blockIdx.x * 128ll + threadIdx.x;
    if (j >= n) return;
    constexpr unsigned PSET = ORDER == 0 ? kPanelsValue : (ORDER == 1 ? kPanelsGrad : kPanelsAll);
    constexpr int NP = popcount10(PSET);
    double px, py, pz;
    point_coords(ps, j, px, py, pz);
    const long long nao = m.nao;
    for (int s = 0; s < m.nshells; ++s) {
        const DevShell sh = m.shells[s];
        const double x = px - sh.o[0], y = py - sh.o[1], z = pz - sh.o[2];
        const double r2 = sq_norm(x, y, z);
        if (r2 <= sh.cutoff_r2) {
            eval_block<ORDER, PSET>(sh, m, x, y, z, r2, [&](int row, const double (&v)[NP]) __attribute__((always_inline)) {
#pragma unroll
                for (int i = 0; i < NP; ++i)
                    out[(sh.ao + row) + nao * (i * n + j)] = v[i];
            });
        } else {
            for (int row = 0; row < sh.count; ++row)
#pragma unroll
                for (int i = 0; i < NP; ++i)
                    out[(sh.ao + row) + nao * (i * n + j)] = 0.0;
        }
    }
}

} // namespace

int max_chunks_bound(const DevShell* shells, int nshells) {
    int count[kMaxL + 1] = {};
    for (int s = 0; s < nshells; ++s)
        ++count[shells[s].l];
    int bound = 1;
    for (int l = 0; l <= kMaxL; ++l) {
        if (!count[l]) continue;
        const int per = kChunkRows / ((l + 1) * (l + 2) / 2);
        bound += (count[l] + per - 1) / per + 1;
    }
    return bound;
}

void launch_masks(const MolView& m, const PointSource& ps, long long npoints, int block_size, long long first_block,
                  int nblk, int order, double threshold, const KeptTables& t, hipStream_t stream) {
    if (nblk <= 0) return;
    const size_t shmem = static_cast<size_t>(m.nao) * sizeof(unsigned long long) +
                         static_cast<size_t>(m.nshells) * 3 * sizeof(int32_t) +
                         static_cast<size_t>(t.max_chunks + 1) * sizeof(int32_t);
    if (shmem > 64 * 1024) invalid("basis too large for the screening kernel (more than about 6000 AOs)");
    const dim3 grid(static_cast<unsigned>(nblk)), block(kMaskThreads);
    switch (order) {
    case 0: mask_kernel<0><<<grid, block, shmem, stream>>>(m, ps, npoints, block_size, first_block, threshold, t); break;
    case 1: mask_kernel<1><<<grid, block, shmem, stream>>>(m, ps, npoints, block_size, first_block, threshold, t); break;
    default: mask_kernel<2><<<grid, block, shmem, stream>>>(m, ps, npoints, block_size, first_block, threshold, t); break;
    }
    KUMO_HIP(hipGetLastError());
}

void launch_ao_eval(const MolView& m, const PointSource& ps, long long n, int order, double* out, hipStream_t stream) {
    if (n <= 0) return;
    const dim3 grid(static_cast<unsigned>((n + 127) / 128)), block(128);
    switch (order) {
    case 0: ao_eval_kernel<0><<<grid, block, 0, stream>>>(m, ps, n, out); break;
    case 1: ao_eval_kernel<1><<<grid, block, 0, stream>>>(m, ps, n, out); break;
    default: ao_eval_kernel<2><<<grid, block, 0, stream>>>(m, ps, n, out); break;
    }
    KUMO_HIP(hipGetLastError());
}

} // namespace kumo
