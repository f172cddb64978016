// Far-field ESP on the GPU, one workgroup per target cluster
// (expand_far_field and local_polynomial, src/electrostatic_potential.cpp:701-844).
//
// far_targets_kernel, per target cluster (centre b):
//   1. every pair gets a state: its source cluster is far enough from b (to
//      be translated), or its error level at |Q - b|^2 (-1 below the first
//      level's distance); the levels are counted;
//   2. one thread finds the cut of the pair budget (levels from the smallest
//      error up, every pair of a level while the budget holds, then the
//      number of pairs of the first level that fits only in part, taken in
//      the reference's pair order);
//   3. the pairs not translated are split, in pair order, into expanded ones
//      (above the cut, or at it within the partial count) and near ones
//      (summed exactly at the points), listed from the two ends of the
//      target's list;
//   4. each wave takes every fourth far source cluster and every fourth
//      expanded pair: the Hermite recursion of 1/|X| (cluster) or of the
//      pair's charge (Boys seeds) runs across the wave in LDS, and each lane
//      accumulates its local coefficients beta (cluster: sum over nu of
//      Q^(N - |beta|)_nu R_(nu + beta); pair: sum over its terms of
//      c_k R_(alpha_k + beta)); the waves' sums are added in a fixed order and
//      scaled by (-1)^|beta| / beta!.
// far_poly_kernel and far_near_kernel, per target cluster: each thread one
// point, nuclear - (local polynomial at r - b + the exact sums of the near
// pairs).
#include "far_field.hpp"

#include "common.hpp"
#include "esp_common.hpp"

namespace kumo {
namespace {
using namespace esp;

constexpr int FT = 256;                           // threads per target workgroup
constexpr int FW = FT / 64;                       // waves
constexpr int NLOC = far_hcount(kFarMultipole);   // local coefficients of a target (|beta| <= N)
constexpr int NLOW = far_hcount(kFarLocal);       // of which a pair expansion reaches (|beta| <= K)
constexpr int RPL = (NLOC + 63) / 64;             // local coefficients per lane
constexpr int kExpandAtCut = 64;                  // state of a pair of the cut level within the budget

// Position of (t, u, v) in the Hermite order (by t + u + v, then t and u descending).
__device__ __attribute__((always_inline)) inline int tuv_index(int t, int u, int v) {
    const int s = t + u + v;
    return far_hcount(s - 1) + (s - t) * (s - t + 1) / 2 + (s - t - u);
}

__device__ __attribute__((always_inline)) inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Recursion recipe of Hermite entry k > 0 (t, u, v): on the first nonzero
// index a, R^n_k = X_a R^(n+1)_i1 + m R^(n+1)_i2 with i1 the entry one lower
// in a, i2 two lower (m = index_a - 1, i2 = 0 when m = 0); packed as i1 |
// i2 << 11 | a << 22 | m << 24.
__device__ __attribute__((always_inline)) inline uint32_t hermite_recipe(int t, int u, int v) {
    int a, m, i1, i2;
    if (t > 0) {
        a = 0, m = t - 1, i1 = tuv_index(t - 1, u, v), i2 = t > 1 ? tuv_index(t - 2, u, v) : 0;
    } else if (u > 0) {
        a = 1, m = u - 1, i1 = tuv_index(t, u - 1, v), i2 = u > 1 ? tuv_index(t, u - 2, v) : 0;
    } else {
        a = 2, m = v - 1, i1 = tuv_index(t, u, v - 1), i2 = v > 1 ? tuv_index(t, u, v - 2) : 0;
    }
    return static_cast<uint32_t>(i1) | static_cast<uint32_t>(i2) << 11 | static_cast<uint32_t>(a) << 22 |
           static_cast<uint32_t>(m) << 24;
}

// R^0_tuv for t + u + v <= top from the seeds R^n_000 (seeds[n], n <= top),
// from n = top down to 0, the entries of one order across the wave's lanes;
// a and b are the wave's two LDS buffers of far_hcount(top) doubles; returns
// the one holding order 0.
__device__ double* wave_hermite(const double* seeds, const double x[3], int top, double* a, double* b,
                                const uint32_t* recipe, int lane) {
    double* cur = a;
    double* prev = b;
    for (int n = top; n >= 0; --n) {
        double* t_ = cur;
        cur = prev;
        prev = t_;
        const int entries = far_hcount(top - n);
        if (lane == 0) cur[0] = seeds[n];
#pragma unroll 4
        for (int k = lane + 1; k < entries; k += 64) {
            const uint32_t r = recipe[k];
            const int ax = (r >> 22) & 3;
            const double xa = ax == 0 ? x[0] : (ax == 1 ? x[1] : x[2]);
            const double m = static_cast<double>(r >> 24);
            cur[k] = xa * prev[r & 2047u] + m * prev[(r >> 11) & 2047u];
        }
        wave_sync();
    }
    return cur;
}

// sum over the entries alpha of degree <= deg of w[alpha] R[alpha + beta]
// (w in the Hermite order): for each degree and t the entries run over v
// ascending, and so do their sums with beta, so the terms come in contiguous
// runs of w and R.
__device__ __attribute__((always_inline)) inline double shifted_dot(const double* __restrict__ w,
                                                                    const double* __restrict__ R, int deg, int tb,
                                                                    int sb, int vb) {
    double s0 = 0.0, s1 = 0.0;
    int k = 0;
    for (int s = 0; s <= deg; ++s) {
        const int sp = s + sb;
        const int dbase = far_hcount(sp - 1);
        for (int t = s; t >= 0; --t) {
            const int d = sp - t - tb;
            const double* r = R + dbase + d * (d + 1) / 2 + vb;
            const double* ww = w + k;
            const int len = s - t + 1;
            int j = 0;
            for (; j + 1 < len; j += 2) {
                s0 += ww[j] * r[j];
                s1 += ww[j + 1] * r[j + 1];
            }
            if (j < len) s0 += ww[j] * r[j];
            k += len;
        }
    }
    return s0 + s1;
}

constexpr int kSeedMax = 2 * kMaxL + kFarLocal; // highest seed order of a pair expansion (L + 6)

// F_0..F_top(T) (top <= kSeedMax) in registers, in boys()'s arithmetic (the
// nearest table point and the 7-term Taylor series below T = 40, the closed
// form and upward recursion above): the loops run over the fixed bound with
// the orders above top left alone, so that the array needs no scratch.
__device__ __attribute__((always_inline)) inline void boys_regs(const double* __restrict__ table, double T, int top,
                                                                double (&v)[kSeedMax + 1]) {
    const double decay = exp(-T);
    if (T >= 40.0) {
        v[0] = 0.5 * sqrt(3.141592653589793 / T);
#pragma unroll
        for (int m = 0; m < kSeedMax; ++m)
            if (m < top) v[m + 1] = ((2.0 * m + 1.0) * v[m] - decay) / (2.0 * T);
        return;
    }
    const int point = static_cast<int>(llround(T / 0.05));
    const double delta = T - point * 0.05;
    const double* row = table + point * kBoysOrders + top;
    double sum = 0.0, factor = 1.0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        sum += row[k] * factor;
        factor *= -delta / (k + 1.0);
    }
    double cur = sum;
#pragma unroll
    for (int m = kSeedMax; m >= 0; --m) {
        if (m <= top) v[m] = cur;
        if (m <= top && m > 0) cur = (2.0 * T * cur + decay) / (2.0 * m - 1.0);
    }
}

__device__ __attribute__((always_inline)) inline double lane_double(double v, int j) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(u), j);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(u >> 32), j);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

__global__ void __launch_bounds__(FT) far_targets_kernel(FarPlanDev plan, const EspPairDev* __restrict__ pairs,
                                                         const double* __restrict__ coef,
                                                         const double* __restrict__ boys_table,
                                                         const FarTargetDev* __restrict__ targets,
                                                         double* __restrict__ local, int32_t* __restrict__ near_list,
                                                         int32_t* __restrict__ nnear, int8_t* __restrict__ work) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int top = plan.max_L + kFarMultipole;
    const int H = far_hcount(top);
    // LDS: the recursion recipes recipe[H], the indices tri[NLOC] of the
    // local coefficients, then per wave two recursion buffers, the seeds
    // (top + 2 doubles) and a pair's Hermite coefficients, then the wave sums
    // red[FW][NLOC], the level counts and the scan words.
    uint32_t* recipe = reinterpret_cast<uint32_t*>(smem);
    uchar4* tri = reinterpret_cast<uchar4*>(smem + ((H * 4 + 15) / 16) * 16);
    double* wbuf = reinterpret_cast<double*>(smem + ((H * 4 + 15) / 16) * 16 + NLOC * 4);
    const int ncoef_max = far_hcount(plan.max_L) < 128 ? 128 : far_hcount(plan.max_L);
    const int wstride = 2 * H + top + 2 + ncoef_max;
    double* buf_a = wbuf + wave * wstride;
    double* buf_b = buf_a + H;
    double* seeds = buf_b + H;
    double* cbuf = seeds + top + 2; // the Hermite coefficients of the pair being expanded
    double* red = wbuf + FW * wstride;
    int* counts = reinterpret_cast<int*>(red + FW * NLOC);
    int* scan = counts + kFarLevels + 4; // cut, partial, -, -, then the wave totals [FW]
    const FarTargetDev tg = targets[blockIdx.x];
    const double b0 = tg.c[0], b1 = tg.c[1], b2 = tg.c[2];
    int8_t* state = work + static_cast<long long>(blockIdx.x) * (plan.npairs + plan.nclusters);
    int8_t* far = state + plan.npairs; // per source cluster: translated (1) or not
    int32_t* nl = near_list + static_cast<long long>(blockIdx.x) * plan.npairs;

    for (int k = tid; k < H; k += FT) {
        int s = 0;
        while (far_hcount(s) <= k)
            ++s;
        int r = k - far_hcount(s - 1), t = s;
        while (r >= s - t + 1) {
            r -= s - t + 1;
            --t;
        }
        if (k > 0) recipe[k] = hermite_recipe(t, s - t - r, r);
        if (k < NLOC)
            tri[k] = uchar4{static_cast<unsigned char>(t), static_cast<unsigned char>(s - t - r),
                            static_cast<unsigned char>(r), 0};
    }
    if (tid < kFarLevels) counts[tid] = 0;
    // 0. the source clusters far enough to be translated, decided once so that
    //    the pairs' states and the translations agree
    for (int ci = tid; ci < plan.nclusters; ci += FT) {
#pragma clang fp contract(off) // the decisions' distances rounded as on the host
        const FarClusterDev& cl = plan.clusters[ci];
        const double cx = cl.c[0] - b0, cy = cl.c[1] - b1, cz = cl.c[2] - b2;
        far[ci] = sqrt((cx * cx + cy * cy) + cz * cz) >= cl.min_distance ? 1 : 0;
    }
    __syncthreads();

    // 1. states: -2 translated with its source cluster, else the error level
    for (int i = tid; i < plan.npairs; i += FT) {
        int lev;
        if (far[plan.cluster_of[i]]) {
            lev = -2;
        } else {
#pragma clang fp contract(off)
            const EspPairDev pr = pairs[i];
            const double dx = pr.qx - b0, dy = pr.qy - b1, dz = pr.qz - b2;
            const double d2 = (dx * dx + dy * dy) + dz * dz;
            const float* lim = plan.level_d2 + static_cast<long long>(i) * kFarLevels;
            lev = -1;
            while (lev + 1 < kFarLevels && d2 >= static_cast<double>(lim[lev + 1]))
                ++lev;
            if (lev >= 0) atomicAdd(&counts[lev], 1);
        }
        state[i] = static_cast<int8_t>(lev);
    }
    __syncthreads();
    // 2. the cut of the pair budget
    if (tid == 0) {
        int cut = -1, partial = 0;
        double used = 0.0;
        for (int j = kFarLevels - 1; j >= 0; --j) {
            const double v = plan.levels[j];
            const double n = static_cast<double>(counts[j]);
            if (used + n * v <= plan.pair_tolerance) {
                used += n * v;
                cut = j - 1;
                continue;
            }
            partial = static_cast<int>(floor((plan.pair_tolerance - used) / v));
            cut = j;
            break;
        }
        scan[0] = cut;
        scan[1] = partial;
    }
    __syncthreads();
    const int cut = scan[0], partial = scan[1];
    // 3. the pairs of the cut level expanded within the budget: the first
    //    `partial` of them in the reference's order, marked kExpandAtCut; then
    //    near pairs from the front of the list, expanded pairs from the back,
    //    both in pair order
    int* wtot = scan + 4; // [FW]
    const auto count_before = [&](bool pred, int& total) __attribute__((always_inline)) {
        const unsigned long long bal = __ballot(pred);
        const int in_wave = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wtot[wave] = __popcll(bal);
        __syncthreads();
        int before = 0;
        total = 0;
        for (int w = 0; w < FW; ++w) {
            if (w < wave) before += wtot[w];
            total += wtot[w];
        }
        __syncthreads();
        return before + in_wave;
    };
    if (cut >= 0 && partial > 0) { // uniform over the workgroup
        int n_cut = 0;
        for (int base = 0; base < plan.npairs && n_cut < partial; base += FT) {
            const int k = base + tid;
            const int i = k < plan.npairs ? plan.ref_order[k] : -1;
            const bool at_cut = i >= 0 && state[i] == cut;
            int tot_cut;
            const int rank = n_cut + count_before(at_cut, tot_cut);
            if (at_cut && rank < partial) state[i] = kExpandAtCut;
            n_cut += tot_cut;
        }
        __syncthreads();
    }
    int n_near = 0, n_exp = 0;
    for (int base = 0; base < plan.npairs; base += FT) {
        const int i = base + tid;
        const int lev = i < plan.npairs ? state[i] : -3;
        int tot_near, tot_exp;
        const bool is_exp = lev >= 0 && lev > cut;
        const bool is_near = !is_exp && lev >= -1;
        const int pn = count_before(is_near, tot_near);
        const int pe = count_before(is_exp, tot_exp);
        if (is_near) nl[n_near + pn] = i;
        if (is_exp) nl[plan.npairs - 1 - (n_exp + pe)] = i;
        n_near += tot_near;
        n_exp += tot_exp;
    }
    if (tid == 0) nnear[blockIdx.x] = n_near;
    __syncthreads(); // the list entries of every wave are written before step 4 reads them

    // 4. the local expansion
    double acc[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r)
        acc[r] = 0.0;
    for (int ci = wave; ci < plan.nclusters; ci += FW) {
        if (!far[ci]) continue; // uniform over the wave
        const FarClusterDev cl = plan.clusters[ci];
        const double x[3] = {cl.c[0] - b0, cl.c[1] - b1, cl.c[2] - b2};
        const double d2 = (x[0] * x[0] + x[1] * x[1]) + x[2] * x[2];
        // seeds (-1)^n (2n - 1)!! / |X|^(2n + 1) (coulomb_derivatives,
        // src/electrostatic_potential.cpp:282-291), to the order the
        // cluster's moments reach: its largest pair order + N (the moments of
        // higher total order are zero, and R^0 of a total order d needs only
        // the seeds to order d, so every term is the one of the full order)
        const int topc = cl.max_L + kFarMultipole;
        if (lane == 0) {
            const double inv = 1.0 / sqrt(d2), inv2 = inv * inv;
            double sd = inv;
            for (int n = 0; n <= topc; ++n) {
                seeds[n] = sd;
                sd *= -(2.0 * n + 1.0) * inv2;
            }
        }
        wave_sync();
        const double* R = wave_hermite(seeds, x, topc, buf_a, buf_b, recipe, lane);
        const double* q0 = plan.moments + static_cast<long long>(ci) * (kFarMultipole + 1) * plan.moment_count;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int be = lane + 64 * r;
            if (be >= NLOC) continue;
            const uchar4 hb = tri[be];
            const int sb = hb.x + hb.y + hb.z;
            const int m = kFarMultipole - sb;
            acc[r] += shifted_dot(q0 + static_cast<long long>(m) * plan.moment_count, R, cl.max_L + m, hb.x, sb,
                                  hb.z);
        }
        wave_sync(); // the buffers are rewritten by the next item
    }
    // Expanded pairs in batches of 64 per wave: each lane loads one pair of
    // the batch and computes its seeds (-2p)^n F_n(p |PC|^2) (hermite_integrals,
    // src/electrostatic_potential.cpp:269-280) in registers; the pairs are
    // then expanded one after another, each pair's seeds and Hermite
    // coefficients staged in LDS, the next pair's coefficients loaded while
    // the current one is expanded.
    const int nexp = n_exp;
    const int ncoef = far_hcount(plan.max_L);
    for (int base = wave * 64; base < nexp; base += FW * 64) {
        const int cnt = min(64, nexp - base);
        const int i = nl[plan.npairs - 1 - (base + min(lane, cnt - 1))];
        const EspPairDev pr = pairs[i];
        const double px = pr.qx - b0, py = pr.qy - b1, pz = pr.qz - b2;
        const int ord = pr.L + kFarLocal;
        double sd[kSeedMax + 1];
        boys_regs(boys_table, pr.p * ((px * px + py * py) + pz * pz), ord, sd);
        {
            double power = 1.0;
#pragma unroll
            for (int n = 1; n <= kSeedMax; ++n) {
                power *= -2.0 * pr.p;
                if (n <= ord) sd[n] *= power;
            }
        }
        // coefficients of the batch's first pair
        const int first0 = __builtin_amdgcn_readlane(pr.first, 0);
        const int nk0 = far_hcount(__builtin_amdgcn_readlane(pr.L, 0));
        double cpre[2] = {lane < nk0 ? coef[first0 + lane] : 0.0, lane + 64 < nk0 ? coef[first0 + lane + 64] : 0.0};
        for (int j = 0; j < cnt; ++j) {
            const int Lj = __builtin_amdgcn_readlane(pr.L, j);
            const int firstj = __builtin_amdgcn_readlane(pr.first, j);
            const int nk = far_hcount(Lj);
            if (lane == j) {
#pragma unroll
                for (int n = 0; n <= kSeedMax; ++n)
                    if (n <= ord) seeds[n] = sd[n];
            }
            cbuf[lane] = cpre[0];
            if (lane + 64 < ncoef) cbuf[lane + 64] = cpre[1];
            for (int k = lane + 128; k < nk; k += 64) // beyond two chunks (L >= 7): straight from memory
                cbuf[k] = coef[firstj + k];
            if (j + 1 < cnt) {
                const int fn = __builtin_amdgcn_readlane(pr.first, j + 1);
                const int nn = far_hcount(__builtin_amdgcn_readlane(pr.L, j + 1));
                cpre[0] = lane < nn ? coef[fn + lane] : 0.0;
                cpre[1] = lane + 64 < nn ? coef[fn + lane + 64] : 0.0;
            }
            wave_sync();
            const double xj[3] = {lane_double(px, j), lane_double(py, j), lane_double(pz, j)};
            const double* R = wave_hermite(seeds, xj, Lj + kFarLocal, buf_a, buf_b, recipe, lane);
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int be = lane + 64 * r;
                if (be >= NLOW) continue;
                const uchar4 hb = tri[be];
                acc[r] += shifted_dot(cbuf, R, Lj, hb.x, hb.x + hb.y + hb.z, hb.z);
            }
            wave_sync();
        }
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int be = lane + 64 * r;
        if (be < NLOC) red[wave * NLOC + be] = acc[r];
    }
    __syncthreads();
    for (int be = tid; be < NLOC; be += FT) {
        double s = 0.0;
        for (int w = 0; w < FW; ++w)
            s += red[w * NLOC + be];
        const uchar4 hb = tri[be];
        double f = ((hb.x + hb.y + hb.z) & 1) ? -1.0 : 1.0;
        for (int j = 2; j <= hb.x; ++j)
            f /= j;
        for (int j = 2; j <= hb.y; ++j)
            f /= j;
        for (int j = 2; j <= hb.z; ++j)
            f /= j;
        local[static_cast<long long>(blockIdx.x) * NLOC + be] = s * f;
    }
}

// Point k of a target: its index in the point source, or -1 where a grid
// tile's point lies outside the launch's range [q0, q1).
__device__ __attribute__((always_inline)) inline long long far_point(const PointSource& ps, const FarTargetDev& tg,
                                                                     const int64_t* idx, int k, long long q0,
                                                                     long long q1) {
    if (ps.points) return idx[tg.first + k];
    const int ey = tg.hi[1] - tg.lo[1], ez = tg.hi[2] - tg.lo[2];
    const int di = k / (ey * ez), r = k - di * ey * ez, dj = r / ez, dk = r - dj * ez;
    const long long q = (static_cast<long long>(tg.lo[0] + di) * ps.ny + (tg.lo[1] + dj)) * ps.nz + (tg.lo[2] + dk);
    return (q >= q0 && q < q1) ? q : -1;
}

// The local polynomial and the nuclear term at the points of each target:
// out = nuclear - sum_beta local_beta (r - b)^beta (the near pairs are
// subtracted by far_near_kernel, one L after another).
__global__ void __launch_bounds__(FT) far_poly_kernel(const EspAtomDev* __restrict__ atoms, int natoms, PointSource ps,
                                                      long long q0, long long q1,
                                                      const FarTargetDev* __restrict__ targets,
                                                      const int64_t* __restrict__ idx,
                                                      const double* __restrict__ local, double* __restrict__ out) {
    const FarTargetDev tg = targets[blockIdx.x];
    const double* loc = local + static_cast<long long>(blockIdx.x) * NLOC;
    for (int k = threadIdx.x; k < tg.count; k += FT) {
        const long long q = far_point(ps, tg, idx, k, q0, q1);
        if (q < 0) continue;
        double x, y, z;
        point_coords(ps, q, x, y, z);
        const double h[3] = {x - tg.c[0], y - tg.c[1], z - tg.c[2]};
        double pw[3][kFarMultipole + 1];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            pw[a][0] = 1.0;
#pragma unroll
            for (int j = 1; j <= kFarMultipole; ++j)
                pw[a][j] = pw[a][j - 1] * h[a];
        }
        double poly = 0.0;
        int be = 0;
#pragma unroll
        for (int s = 0; s <= kFarMultipole; ++s)
#pragma unroll
            for (int t = s; t >= 0; --t)
#pragma unroll
                for (int u = s - t; u >= 0; --u, ++be)
                    poly += loc[be] * pw[0][t] * pw[1][u] * pw[2][s - t - u];
        double nuclear = 0.0;
        for (int a = 0; a < natoms; ++a) {
            const EspAtomDev at = atoms[a];
            const double dx = at.x - x, dy = at.y - y, dz = at.z - z;
            const double d = sqrt((dx * dx + dy * dy) + dz * dz);
            if (at.charge == 0.0 || d < 1e-8) continue;
            nuclear += at.charge / d;
        }
        out[q] = nuclear - poly;
    }
}

// The near pairs of order L (pair indices [lo, hi), a contiguous part of
// each target's sorted near list), subtracted at the target's points, 64
// points at a time with the pairs split into 4 slices over the threads,
// summed in slice order.
template <int L>
__global__ void __launch_bounds__(FT) far_near_kernel(const EspPairDev* __restrict__ pairs,
                                                      const double* __restrict__ coef,
                                                      const double* __restrict__ boys_table, PointSource ps,
                                                      long long q0, long long q1,
                                                      const FarTargetDev* __restrict__ targets,
                                                      const int64_t* __restrict__ idx,
                                                      const int32_t* __restrict__ near_list,
                                                      const int32_t* __restrict__ nnear, int npairs, int lo, int hi,
                                                      double* __restrict__ out) {
    __shared__ double part[FT];
    const FarTargetDev tg = targets[blockIdx.x];
    const int32_t* nl = near_list + static_cast<long long>(blockIdx.x) * npairs;
    const int nn = nnear[blockIdx.x];
    const auto lower = [&](int key) __attribute__((always_inline)) {
        int a = 0, b = nn;
        while (a < b) {
            const int mid = (a + b) >> 1;
            if (nl[mid] < key) a = mid + 1;
            else b = mid;
        }
        return a;
    };
    const int e0 = lower(lo), e1 = lower(hi);
    if (e0 >= e1) return; // uniform over the workgroup
    // 64 points per pass, the pairs in 4 slices: the same summation order for
    // a point whatever else its target holds (a range of listed points holds
    // part of a cube)
    constexpr int span = 64, slices = FT / span;
    const int slice = threadIdx.x / span;
    for (int k0 = 0; k0 < tg.count; k0 += span) {
        const int k = k0 + static_cast<int>(threadIdx.x) % span;
        long long q = -1;
        double sum = 0.0;
        if (k < tg.count) q = far_point(ps, tg, idx, k, q0, q1);
        if (q >= 0) {
            double x, y, z;
            point_coords(ps, q, x, y, z);
            for (int e = e0 + slice; e < e1; e += slices)
                sum += pair_sum<L>(pairs[nl[e]], coef, boys_table, x, y, z);
        }
        part[threadIdx.x] = sum;
        __syncthreads();
        if (slice == 0) {
#pragma unroll
            for (int sl = 1; sl < slices; ++sl)
                sum += part[threadIdx.x + sl * span];
        }
        __syncthreads();
        if (slice == 0 && q >= 0) out[q] -= sum;
    }
}

} // namespace

void launch_far_targets(const FarPlanDev& plan, const EspPairDev* pairs, const double* coef, const double* boys_table,
                        const FarTargetDev* targets, int ntargets, double* local, int32_t* near, int32_t* nnear,
                        int8_t* work, hipStream_t stream) {
    if (ntargets <= 0) return;
    const int top = plan.max_L + kFarMultipole;
    const int H = far_hcount(top);
    if (H > 2048) invalid("far-field expansion order too high"); // recipe indices take 11 bits
    const int ncoef_max = far_hcount(plan.max_L) < 128 ? 128 : far_hcount(plan.max_L);
    const size_t lds = ((H * 4 + 15) / 16) * 16 + NLOC * 4 +
                       sizeof(double) * (FW * (2 * H + top + 2 + ncoef_max) + FW * NLOC) +
                       sizeof(int) * (kFarLevels + 4 + 4 + FW);
    if (lds > 160 * 1024) invalid("far-field expansion needs more LDS than a workgroup has");
    // (set on every launch: the attribute belongs to the current device)
    KUMO_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&far_targets_kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    far_targets_kernel<<<ntargets, FT, lds, stream>>>(plan, pairs, coef, boys_table, targets, local, near, nnear,
                                                      work);
    KUMO_HIP(hipGetLastError());
}

void launch_far_points(const EspPairDev* pairs, const std::vector<EspGroup>& groups, const double* coef,
                       const double* boys_table, const EspAtomDev* atoms, int natoms, const PointSource& ps,
                       long long q0, long long q1, const FarTargetDev* targets, int ntargets, const int64_t* idx,
                       const double* local, const int32_t* near, const int32_t* nnear, int npairs, double* out,
                       hipStream_t stream) {
    if (ntargets <= 0) return;
    far_poly_kernel<<<ntargets, FT, 0, stream>>>(atoms, natoms, ps, q0, q1, targets, idx, local, out);
    KUMO_HIP(hipGetLastError());
    for (const EspGroup& g : groups) {
        if (g.last <= g.first) continue;
#define KUMO_FAR_NEAR(LL)                                                                                             \
    far_near_kernel<LL><<<ntargets, FT, 0, stream>>>(pairs, coef, boys_table, ps, q0, q1, targets, idx, near, nnear,   \
                                                     npairs, g.first, g.last, out)
        switch (g.L) {
        case 0: KUMO_FAR_NEAR(0); break;
        case 1: KUMO_FAR_NEAR(1); break;
        case 2: KUMO_FAR_NEAR(2); break;
        case 3: KUMO_FAR_NEAR(3); break;
        case 4: KUMO_FAR_NEAR(4); break;
        case 5: KUMO_FAR_NEAR(5); break;
        case 6: KUMO_FAR_NEAR(6); break;
        case 7: KUMO_FAR_NEAR(7); break;
        case 8: KUMO_FAR_NEAR(8); break;
        case 9: KUMO_FAR_NEAR(9); break;
        case 10: KUMO_FAR_NEAR(10); break;
        default: invalid("invalid electrostatic potential pair group");
        }
#undef KUMO_FAR_NEAR
        KUMO_HIP(hipGetLastError());
    }
}

} // namespace kumo
