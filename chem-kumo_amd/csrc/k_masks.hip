// Screening masks (src/ao_screening.hpp:35-50 as applied per point block in
// src/density.cpp:454-467) and AO panel evaluation.
//
// significant_aos keeps AO r of a block of points when the largest |value| of
// its screened panels (value; + gradient; + Hessian) over the block's points
// exceeds the threshold. One workgroup per screening block decides every AO
// block in four phases, and the decisions equal those of the reference
// arithmetic (src/basis_evaluator.cpp:512-643, no fused multiply-adds; this
// file is built with -ffp-contract=off, see ao_eval.hpp):
//   A. lanes = AO blocks (in descending l): the block's point nearest to the
//      AO block's centre gives the smallest distance. Beyond the cutoff, or
//      beyond the radius where the rigorous bound of
//      src/basis_evaluator.cpp:98-127 falls to the threshold, every AO is
//      dropped. Otherwise the AOs are evaluated at that point with the fast
//      factored arithmetic; when each AO's largest panel there exceeds the
//      threshold by more than a relative 1e-6 (far beyond the few-ulp
//      difference of the two arithmetics), every AO is kept (the maximum
//      over the block is at least that value);
//   B. lanes = points: the remaining AO blocks are evaluated at every point
//      with the fast arithmetic and each AO's maximum is reduced (wave
//      shuffles, then a 64-bit LDS atomic max on the IEEE bit pattern,
//      monotone for non-negative doubles);
//   C. AO blocks with an AO whose fast maximum lies within a relative 1e-6 of
//      the threshold are evaluated again with the reference arithmetic, which
//      decides them;
//   D. the kept AO list, the kept AO blocks (in AO order for the AO-P path)
//      and the AO->MO chunk plan (entries by descending l, packed next-fit)
//      are written for the contraction kernels.
#include "kernels.hpp"

#include "common.hpp"

namespace kumo {
namespace {

constexpr int kMaskThreads = 128; // one lane per point of a default 128-point block
constexpr double kWindow = 1e-6;  // relative band around the threshold decided exactly

enum : int32_t { kDrop = 0, kKeepAll = 1, kFull = 2, kExact = 3 };

// Calls f.template operator()<C>() for each derivative component C of PSET.
template <unsigned PSET, int I = 0, typename F> __device__ __attribute__((always_inline)) inline void for_panels(F&& f) {
    if constexpr (I < popcount10(PSET)) {
        f.template operator()<panel_component(PSET, I)>();
        for_panels<PSET, I + 1>(f);
    }
}

// Fast largest |panel| of each AO of block sh (angular momentum L) at one point.
template <int L, int ORDER, unsigned PSET>
__device__ __attribute__((always_inline)) inline void fast_block_max(const DevShell& sh, const MolView& m, double x,
                                                                   double y, double z, const double (&R)[3],
                                                                   double (&amax)[kMaxCart]) {
#pragma clang fp contract(fast)
    constexpr int NC = (L + 1) * (L + 2) / 2;
    const bool pure = L >= 2 && sh.pure;
    const double* tg = m.tg + L * kMaxPure * kMaxCart;
    const double* gk = m.gk + L * kMaxCart;
    for_panels<PSET>([&]<int C>() __attribute__((always_inline)) {
        double cv[NC];
        cart_components<L, ORDER, (1u << C)>(x, y, z, R, [&](int k, const double (&v)[1]) __attribute__((always_inline)) {
            cv[k] = v[0];
        });
        if (pure) {
#pragma unroll
            for (int r = 0; r < 2 * L + 1; ++r) {
                double a = 0.0;
#pragma unroll
                for (int k = 0; k < NC; ++k)
                    a += tg[r * kMaxCart + k] * cv[k];
                amax[r] = fmax(amax[r], fabs(a));
            }
        } else {
#pragma unroll
            for (int k = 0; k < NC; ++k)
                amax[k] = fmax(amax[k], fabs(gk[k] * cv[k]));
        }
    });
}

// Fast smallest (over the block's AOs) largest |panel| at one point, or any
// value above `pass` not above it (the caller only asks whether it exceeds
// pass): the value panel alone is tried first, since an AO whose value
// already exceeds pass exceeds it over all panels. Lanes take AO blocks in l
// order, so a wave mostly shares l: l <= 3 uses the compile-time evaluation,
// higher l a runtime loop.
template <int ORDER, unsigned PSET>
__device__ __attribute__((always_inline)) inline double probe_min_max(const DevShell& sh, const MolView& m, double x,
                                                                     double y, double z, double r2, double pass) {
#pragma clang fp contract(fast)
    constexpr int NP = popcount10(PSET);
    double R[3];
    radial_sums<ORDER>(sh, m, r2, R);
    if (sh.l <= 3) {
        double mn = 1.0e300;
        const auto probe = [&]<unsigned PS>() __attribute__((always_inline)) {
            double amax[kMaxCart];
#pragma unroll
            for (int r = 0; r < kMaxCart; ++r)
                amax[r] = 0.0;
            switch (sh.l) {
            case 0: fast_block_max<0, ORDER, PS>(sh, m, x, y, z, R, amax); break;
            case 1: fast_block_max<1, ORDER, PS>(sh, m, x, y, z, R, amax); break;
            case 2: fast_block_max<2, ORDER, PS>(sh, m, x, y, z, R, amax); break;
            default: fast_block_max<3, ORDER, PS>(sh, m, x, y, z, R, amax); break;
            }
            mn = 1.0e300;
#pragma unroll
            for (int r = 0; r < 10; ++r)
                if (r < sh.count) mn = fmin(mn, amax[r]);
        };
        probe.template operator()<kPanelsValue>();
        if (PSET != kPanelsValue && !(mn > pass)) probe.template operator()<PSET>();
        return mn;
    }
    const int L = sh.l, nc = sh.ncart;
    double mn = 1.0e300;
    if (sh.pure) {
        const double* tg = m.tg + L * kMaxPure * kMaxCart;
        for (int r = 0; r < sh.count; ++r) {
            double v[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i)
                v[i] = 0.0;
            for (int k = 0; k < nc; ++k) {
                const double w = tg[r * kMaxCart + k];
                if (w == 0.0) continue;
                double cv[NP];
                cart_component_rt<ORDER, PSET>(L, k, x, y, z, R, cv);
#pragma unroll
                for (int i = 0; i < NP; ++i)
                    v[i] += w * cv[i];
            }
            double mx = 0.0;
#pragma unroll
            for (int i = 0; i < NP; ++i)
                mx = fmax(mx, fabs(v[i]));
            mn = fmin(mn, mx);
        }
    } else {
        const double* gk = m.gk + L * kMaxCart;
        for (int k = 0; k < nc; ++k) {
            double cv[NP];
            cart_component_rt<ORDER, PSET>(L, k, x, y, z, R, cv);
            double mx = 0.0;
#pragma unroll
            for (int i = 0; i < NP; ++i)
                mx = fmax(mx, fabs(gk[k] * cv[i]));
            mn = fmin(mn, mx);
        }
    }
    return mn;
}

// Smallest squared distance from c to the block's points (cached in spts),
// and the first point reaching it. For a grid the block's points lie on at
// most a few z-line segments, along each of which the squared distance is a
// convex quadratic in k: only the points within two steps of its clamped
// minimiser are evaluated (with the point-by-point arithmetic, so the value
// is the minimum the full scan finds). Listed points are scanned.
__device__ __attribute__((always_inline)) inline void nearest_point(const PointSource& ps, long long start, int count,
                                                                  const double* spts, const double (&c)[3],
                                                                  double& best, int& bj) {
    best = 1.0e300;
    bj = 0;
    const auto probe = [&](int jj) __attribute__((always_inline)) {
        const double r2 = sq_norm(spts[4 * jj] - c[0], spts[4 * jj + 1] - c[1], spts[4 * jj + 2] - c[2]);
        if (r2 < best || (r2 == best && jj < bj)) {
            best = r2;
            bj = jj;
        }
    };
    if (ps.points) {
        for (int jj = 0; jj < count; ++jj)
            probe(jj);
        return;
    }
    const double* s = ps.steps;
    const double dd = s[6] * s[6] + s[7] * s[7] + s[8] * s[8];
    const long long end = start + count;
    for (long long q = start; q < end;) {
        const long long line = q / ps.nz, k0 = q % ps.nz;
        const long long k1 = min(ps.nz - 1, k0 + (end - q) - 1);
        const double fi = static_cast<double>(line / ps.ny), fj = static_cast<double>(line % ps.ny);
        double a[3];
        for (int e = 0; e < 3; ++e)
            a[e] = ps.origin[e] + s[e] * fi + s[3 + e] * fj - c[e];
        double ks = -(a[0] * s[6] + a[1] * s[7] + a[2] * s[8]) / dd;
        ks = fmin(fmax(ks, static_cast<double>(k0)), static_cast<double>(k1));
        const long long kc = llrint(ks);
        for (long long kk = kc - 2; kk <= kc + 2; ++kk)
            if (kk >= k0 && kk <= k1) probe(static_cast<int>(line * ps.nz + kk - start));
        q = (line + 1) * ps.nz;
    }
}

// Exclusive prefix sum of v over the workgroup; total receives the sum.
__device__ __attribute__((always_inline)) inline int block_scan(int v, int* tmp, int& total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    const int before = w ? tmp[0] : 0;
    total = tmp[0] + tmp[1];
    __syncthreads();
    return before + x - v;
}

// At least 3 waves per SIMD: the Hessian-order kernel then fits 168 VGPRs
// (43 spilled to scratch) instead of 217 at 2 waves, and runs 19 % faster on
// a 640-AO molecule's 160^3 cube; the lower orders fit in 3 waves anyway.
template <int ORDER>
__global__ void __launch_bounds__(kMaskThreads) __attribute__((amdgpu_waves_per_eu(3)))
    mask_kernel(MolView m, PointSource ps, long long npoints, int block_size, long long first_block, double threshold,
                const double* screen_r2, KeptTables t) {
    constexpr unsigned PSET = ORDER == 0 ? kPanelsValue : (ORDER == 1 ? kPanelsGrad : kPanelsAll);
    constexpr int NP = popcount10(PSET);
    // LDS: smax[nao] (u64), pts[kMaskThreads][4] (double), then per AO block
    // bits, state, entry, entry slot (int32), then rows used per chunk, then
    // the lists of the AO blocks left to phases B and C (int32, nshells each).
    extern __shared__ unsigned long long smax[];
    const int nao = m.nao, nsh = m.nshells;
    double* spts = reinterpret_cast<double*>(smax + nao);
    uint32_t* sbits = reinterpret_cast<uint32_t*>(spts + 4 * kMaskThreads);
    int32_t* sstate = reinterpret_cast<int32_t*>(sbits + nsh);
    int32_t* sent = sstate + nsh;
    int32_t* sslot = sent + nsh;
    int32_t* sused = sslot + nsh;
    int32_t* lfull = sused + t.max_chunks + 1;
    int32_t* lexact = lfull + nsh;
    __shared__ int scan_tmp[2];
    __shared__ int packed[2];
    __shared__ int nfull, nexact;

    const int tid = threadIdx.x, lane = tid & 63;
    const long long blk = first_block + blockIdx.x;
    const long long start = blk * block_size;
    const int count = static_cast<int>(min(static_cast<long long>(block_size), npoints - start));
    const bool screen = threshold != 0.0 && count > 0;
    for (int a = tid; a < nao; a += kMaskThreads)
        smax[a] = 0ull;
    if (tid == 0) nfull = nexact = 0;
    // AO blocks whose maxima phases B and C evaluate (in any order: the
    // maxima are order-independent)
    const auto enlist = [&](int s, int st) __attribute__((always_inline)) {
        if (st == kFull) lfull[atomicAdd(&nfull, 1)] = s;
        else if (st == kExact) lexact[atomicAdd(&nexact, 1)] = s;
    };
    const bool cached = count <= kMaskThreads; // points held in LDS
    if (cached && tid < count) {
        double px, py, pz;
        point_coords(ps, start + tid, px, py, pz);
        spts[4 * tid] = px;
        spts[4 * tid + 1] = py;
        spts[4 * tid + 2] = pz;
    }
    __syncthreads();

    if (screen && !cached) {
        // Blocks of more than kMaskThreads points skip the classification.
        for (int s = tid; s < nsh; s += kMaskThreads) {
            const int st = m.shells[s].l > 3 ? kExact : kFull;
            sstate[s] = st;
            enlist(s, st);
        }
        __syncthreads();
    }
    if (screen && cached) {
        // A. Classification, one lane per AO block.
        for (int q = tid; q < nsh; q += kMaskThreads) {
            const int s = m.lorder[q];
            const DevShell& sh = m.shells[s];
            double best;
            int bj;
            const double c[3] = {sh.o[0], sh.o[1], sh.o[2]};
            nearest_point(ps, start, count, spts, c, best, bj);
            int st;
            if (!(best <= sh.cutoff_r2) || best > screen_r2[s]) {
                st = kDrop;
            } else {
                const double x = spts[4 * bj] - sh.o[0], y = spts[4 * bj + 1] - sh.o[1], z = spts[4 * bj + 2] - sh.o[2];
                const double pass = threshold * (1.0 + kWindow);
                const double mn = probe_min_max<ORDER, PSET>(sh, m, x, y, z, best, pass);
                st = mn > pass ? kKeepAll : (sh.l > 3 ? kExact : kFull);
            }
            sstate[s] = st;
            enlist(s, st);
        }
        __syncthreads();
    }
    if (screen) {
        // B. Fast maxima over the points of the undecided blocks.
        const int nf = nfull;
        for (int p0 = 0; p0 < count; p0 += kMaskThreads) {
        const bool active = p0 + tid < count;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (active) {
            if (cached) {
                px = spts[4 * tid];
                py = spts[4 * tid + 1];
                pz = spts[4 * tid + 2];
            } else {
                point_coords(ps, start + p0 + tid, px, py, pz);
            }
        }
        for (int u = 0; u < nf; ++u) {
            const int s = lfull[u];
            const DevShell& sh = m.shells[s];
            const double x = px - sh.o[0], y = py - sh.o[1], z = pz - sh.o[2];
            const double r2 = sq_norm(x, y, z);
            const bool inside = active && r2 <= sh.cutoff_r2;
            if (!__any(inside)) continue;
            double amax[kMaxCart];
#pragma unroll
            for (int r = 0; r < kMaxCart; ++r)
                amax[r] = 0.0;
            if (inside) {
                double R[3];
                radial_sums<ORDER>(sh, m, r2, R);
                switch (sh.l) {
                case 0: fast_block_max<0, ORDER, PSET>(sh, m, x, y, z, R, amax); break;
                case 1: fast_block_max<1, ORDER, PSET>(sh, m, x, y, z, R, amax); break;
                case 2: fast_block_max<2, ORDER, PSET>(sh, m, x, y, z, R, amax); break;
                default: fast_block_max<3, ORDER, PSET>(sh, m, x, y, z, R, amax); break;
                }
            }
#pragma unroll
            for (int r = 0; r < 2 * 3 + 1 + 3; ++r) { // at most 10 rows for l <= 3
                if (r < sh.count) {
                    const double v = wave_max(amax[r]);
                    if (lane == 0 && v > 0.0)
                        atomicMax(&smax[sh.ao + r], static_cast<unsigned long long>(__double_as_longlong(v)));
                }
            }
        }
        }
        __syncthreads();
        // C. Blocks with an AO near the threshold are decided by the reference
        //    arithmetic.
        for (int u = tid; u < nf; u += kMaskThreads) {
            const int s = lfull[u];
            const DevShell& sh = m.shells[s];
            bool near = false;
            for (int r = 0; r < sh.count; ++r) {
                const double v = __longlong_as_double(static_cast<long long>(smax[sh.ao + r]));
                near = near || fabs(v - threshold) <= kWindow * threshold;
            }
            if (near) {
                sstate[s] = kExact;
                enlist(s, kExact);
            }
        }
        __syncthreads();
        const int ne = nexact;
        for (int u = tid; u < ne; u += kMaskThreads) {
            const DevShell& sh = m.shells[lexact[u]];
            for (int r = 0; r < sh.count; ++r)
                smax[sh.ao + r] = 0ull;
        }
        __syncthreads();
        for (int p0 = 0; p0 < count; p0 += kMaskThreads) {
        const bool active = p0 + tid < count;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (active) {
            if (cached) {
                px = spts[4 * tid];
                py = spts[4 * tid + 1];
                pz = spts[4 * tid + 2];
            } else {
                point_coords(ps, start + p0 + tid, px, py, pz);
            }
        }
        for (int u = 0; u < ne; ++u) {
            const int s = lexact[u];
            const DevShell& sh = m.shells[s];
            const double x = px - sh.o[0], y = py - sh.o[1], z = pz - sh.o[2];
            const double r2 = sq_norm(x, y, z);
            const bool inside = active && r2 <= sh.cutoff_r2;
            if (!__any(inside)) continue;
            double amax[kMaxCart];
#pragma unroll
            for (int r = 0; r < kMaxCart; ++r)
                amax[r] = 0.0;
            if (inside) {
                // (rarely reached: runtime-l code keeps this path's registers out of the budget)
                eval_block_rt<ORDER, PSET>(sh, m, x, y, z, r2, [&](int row, const double (&v)[NP]) __attribute__((always_inline)) {
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
                    if (lane == 0 && v > 0.0)
                        atomicMax(&smax[sh.ao + r], static_cast<unsigned long long>(__double_as_longlong(v)));
                }
            }
        }
        }
        __syncthreads();
    }

    // D. Kept bits of every AO block.
    for (int s = tid; s < nsh; s += kMaskThreads) {
        const DevShell& sh = m.shells[s];
        const uint32_t all = sh.count >= 32 ? 0xffffffffu : ((1u << sh.count) - 1u);
        uint32_t bits = all;
        if (screen) {
            const int st = sstate[s];
            if (st == kDrop) {
                bits = 0;
            } else if (st != kKeepAll) {
                bits = 0;
                for (int r = 0; r < sh.count; ++r)
                    if (__longlong_as_double(static_cast<long long>(smax[sh.ao + r])) > threshold) bits |= 1u << r;
            }
        }
        sbits[s] = bits;
    }
    __syncthreads();
    if (!t.lists) {
        // The AO->MO plan needs only the kept AO count.
        if (tid == 0) packed[0] = 0;
        __syncthreads();
        int cnt = 0;
        for (int s = tid; s < nsh; s += kMaskThreads)
            cnt += __popc(sbits[s]);
        if (cnt) atomicAdd(&packed[0], cnt);
        __syncthreads();
        if (tid == 0) t.nkept[blockIdx.x] = packed[0];
    } else {
        // Flags, the ascending kept AO list and the kept blocks in AO order.
        uint8_t* kept = t.kept + static_cast<long long>(blockIdx.x) * nao;
        int32_t* list = t.kept_list + static_cast<long long>(blockIdx.x) * nao;
        int32_t* ks = t.kept_shell + static_cast<long long>(blockIdx.x) * nsh;
        uint32_t* kb = t.kept_bits + static_cast<long long>(blockIdx.x) * nsh;
        int aos = 0, blocks = 0;
        for (int s0 = 0; s0 < nsh; s0 += kMaskThreads) {
            const int s = s0 + tid;
            const uint32_t bits = s < nsh ? sbits[s] : 0u;
            int tot_a, tot_b;
            const int pa = aos + block_scan(__popc(bits), scan_tmp, tot_a);
            const int pb = blocks + block_scan(bits ? 1 : 0, scan_tmp, tot_b);
            if (s < nsh) {
                const DevShell& sh = m.shells[s];
                int w = pa;
                for (int r = 0; r < sh.count; ++r) {
                    const bool k = (bits >> r) & 1u;
                    kept[sh.ao + r] = k ? 1 : 0;
                    if (k) list[w++] = sh.ao + r;
                }
                if (bits) {
                    ks[pb] = s;
                    kb[pb] = bits;
                }
            }
            aos += tot_a;
            blocks += tot_b;
        }
        if (tid == 0) {
            t.nkept[blockIdx.x] = aos;
            t.nkshell[blockIdx.x] = blocks;
        }
    }
    // AO->MO entries: kept blocks in descending l.
    int nent = 0;
    for (int q0 = 0; q0 < nsh; q0 += kMaskThreads) {
        const int q = q0 + tid;
        const int s = q < nsh ? m.lorder[q] : 0;
        const uint32_t bits = q < nsh ? sbits[s] : 0u;
        int tot;
        const int pos = nent + block_scan(bits ? 1 : 0, scan_tmp, tot);
        if (bits) {
            const DevShell& sh = m.shells[s];
            const uint32_t all = sh.count >= 32 ? 0xffffffffu : ((1u << sh.count) - 1u);
            const int mode = bits == all ? kEntryCart : (sh.pure ? kEntryPure : kEntryCartMasked);
            sent[pos] = s | (mode << 24);
            sslot[pos] = sh.ncart; // rows (Cartesian components, also for folded pure blocks), packed below
        }
        nent += tot;
    }
    __syncthreads();
    // Next-fit packing of the entries into chunks of at most kChunkRows rows
    // (each entry has at most kChunkRows rows). With S the exclusive prefix
    // sum of the entries' rows, a chunk starting at entry q ends before the
    // first entry i - 1 with S[i] > S[q] + kChunkRows: the chunk starts are
    // the orbit of 0 under that map, walked by one lane; every entry then
    // finds its chunk by bisection. (smax, lfull and lexact are free here.)
    int32_t* cb = t.chunk_begin + static_cast<long long>(blockIdx.x) * (t.max_chunks + 1);
    int32_t* spre = lfull;                                  // S[q], q < nent
    int32_t* snext = lexact;                                // next chunk start after a chunk starting at q
    int32_t* cstart = reinterpret_cast<int32_t*>(smax);     // chunk starts
    __shared__ int stotal;
    {
        int base = 0;
        for (int q0 = 0; q0 < nent; q0 += kMaskThreads) {
            const int q = q0 + tid;
            const int rows = q < nent ? sslot[q] : 0;
            int tot;
            const int pre = base + block_scan(rows, scan_tmp, tot);
            if (q < nent) spre[q] = pre;
            base += tot;
        }
        if (tid == 0) stotal = base;
        __syncthreads();
        const int total = stotal;
        for (int q = tid; q < nent; q += kMaskThreads) {
            const int v = spre[q] + kChunkRows;
            // first i in [q + 1, nent] with S[i] > v (S[nent] = total)
            int lo = q + 1, hi = nent;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (spre[mid] > v) hi = mid;
                else lo = mid + 1;
            }
            snext[q] = (lo == nent && !(total > v)) ? nent : lo - 1;
        }
        __syncthreads();
        if (tid == 0) {
            int nc = 0, q = 0;
            while (q < nent) {
                if (nc == t.max_chunks) break; // not reached: max_chunks bounds every block
                cstart[nc] = q;
                cb[nc] = q;
                const int nq = snext[q];
                sused[nc] = (nq < nent ? spre[nq] : total) - spre[q];
                ++nc;
                q = nq;
            }
            cb[nc] = q;
            t.nchunks[blockIdx.x] = nc;
            packed[0] = nc;
            packed[1] = q;
        }
        __syncthreads();
        const int nc = packed[0], qend = packed[1];
        for (int q = tid; q < qend; q += kMaskThreads) {
            int lo = 0, hi = nc - 1; // last chunk start <= q
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (cstart[mid] <= q) lo = mid;
                else hi = mid - 1;
            }
            sslot[q] = (lo << 16) | ((q - cstart[lo]) << 8) | (spre[q] - spre[cstart[lo]]);
        }
        __syncthreads();
    }
    const int nc = packed[0];
    nent = packed[1];
    // Fold workspace rows of the partially kept pure entries: the exclusive
    // prefix sum of their Cartesian components, in entry order; the first
    // t.fold_per_block go to the workspace (snext is free here).
    int32_t* sfold = snext;
    if (t.fold_per_block > 0) {
        int base = 0;
        for (int q0 = 0; q0 < nent; q0 += kMaskThreads) {
            const int q = q0 + tid;
            const int rows = q < nent && (sent[q] >> 24) == kEntryPure ? m.shells[sent[q] & 0xffffff].ncart : 0;
            int tot;
            const int pre = base + block_scan(rows, scan_tmp, tot);
            if (q < nent) sfold[q] = pre;
            base += tot;
        }
        if (tid == 0) t.fold_n[blockIdx.x] = min(base, t.fold_per_block);
        __syncthreads();
    }
    int32_t* en = t.ent + static_cast<long long>(blockIdx.x) * nsh;
    uint32_t* eb = t.ent_bits + static_cast<long long>(blockIdx.x) * nsh;
    uint8_t* er = t.ent_row + static_cast<long long>(blockIdx.x) * nsh;
    int32_t* crows = t.chunk_rows + static_cast<long long>(blockIdx.x) * t.max_chunks * kChunkRows * 2;
    for (int q = tid; q < nent; q += kMaskThreads) {
        const int e = sent[q];
        const int s = e & 0xffffff, mode = e >> 24;
        const int c = sslot[q] >> 16, row0 = sslot[q] & 0xff;
        const DevShell& sh = m.shells[s];
        const uint32_t bits = sbits[s];
        en[q] = e;
        eb[q] = bits;
        er[q] = static_cast<uint8_t>(row0);
        // Every entry contributes its Cartesian components: rows of dropped
        // AOs of a Cartesian block get zero coefficients (-1); the rows of a
        // partially kept pure block are folded: coefficient row k is the sum
        // over its kept pure AOs r of T[r][k] g[k] C[r], formed by
        // launch_fold into the fold workspace (code kFoldedWs - workspace
        // row) or, past its rows, when the contraction stages the chunk
        // (code kFoldedRow - (entry << 5 | k)).
        if (c < t.max_chunks)
            for (int i = 0; i < sh.ncart; ++i) {
                int32_t* cr = crows + 2 * (c * kChunkRows + row0 + i);
                if (mode == kEntryPure) {
                    const int code = kFoldedRow - ((q << 5) | i);
                    const int j = t.fold_per_block > 0 ? sfold[q] + i : t.fold_per_block;
                    if (j < t.fold_per_block) {
                        const int w = static_cast<int>(blockIdx.x) * t.fold_per_block + j;
                        t.fold_desc[w] = code;
                        cr[0] = kFoldedWs - w;
                    } else {
                        cr[0] = code;
                    }
                } else {
                    cr[0] = (mode == kEntryCart || ((bits >> i) & 1u)) ? sh.crow + i : -1;
                }
                // Phi row: block << 15 | Cartesian mode << 13 | Cartesian powers (3 bits each)
                const int pm = mode == kEntryPure ? kEntryCartMasked : mode;
                cr[1] = (s << 15) | (pm << 13) | (kPow[sh.l][i][0] << 6) | (kPow[sh.l][i][1] << 3) | kPow[sh.l][i][2];
            }
    }
    for (int item = tid; item < nc * kChunkRows; item += kMaskThreads) {
        const int c = item / kChunkRows, r = item % kChunkRows;
        if (r >= sused[c]) {
            crows[2 * item] = -1;
            crows[2 * item + 1] = -1;
        }
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
                  int nblk, int order, double threshold, const double* screen_r2, const KeptTables& t,
                  hipStream_t stream) {
    if (nblk <= 0) return;
    const size_t shmem = static_cast<size_t>(m.nao) * sizeof(unsigned long long) + 4 * kMaskThreads * sizeof(double) +
                         static_cast<size_t>(m.nshells) * 6 * sizeof(int32_t) +
                         static_cast<size_t>(t.max_chunks + 1) * sizeof(int32_t);
    if (shmem > 64 * 1024) invalid("basis too large for the screening kernel (more than about 6000 AOs)");
    const dim3 grid(static_cast<unsigned>(nblk)), block(kMaskThreads);
    switch (order) {
    case 0: mask_kernel<0><<<grid, block, shmem, stream>>>(m, ps, npoints, block_size, first_block, threshold, screen_r2, t); break;
    case 1: mask_kernel<1><<<grid, block, shmem, stream>>>(m, ps, npoints, block_size, first_block, threshold, screen_r2, t); break;
    default: mask_kernel<2><<<grid, block, shmem, stream>>>(m, ps, npoints, block_size, first_block, threshold, screen_r2, t); break;
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
