// Per-point combination of channel moments into density components
// (combine, src/density.cpp:50-60) and the derived fields of store_block
// (src/evaluator.cpp:127-353), in FP64. One thread per point; every output is
// a coalesced FP64 stream.
#include "kernels.hpp"
#include "plan.hpp"

#include "common.hpp"

#include <math.h>

namespace kumo {
namespace {


struct DDv {
    double v, g[3], t[3], h[6], lap, tau;
};

__device__ inline DDv load_channel(const double* m, long long n, long long j) {
    DDv d;
    d.v = m[M_VALUE * n + j];
    for (int a = 0; a < 3; ++a) {
        d.g[a] = m[(M_GX + a) * n + j];
        d.t[a] = m[(M_TX + a) * n + j];
    }
    for (int k = 0; k < 6; ++k)
        d.h[k] = m[(M_HXX + k) * n + j];
    d.tau = 0.0;
    for (int a = 0; a < 3; ++a)
        d.tau += d.t[a];
    d.lap = d.h[0] + d.h[3] + d.h[5];
    return d;
}

__device__ inline DDv combine(const DDv& a, const DDv& b, double sign, double scale) {
    DDv d;
    d.v = scale * (a.v + sign * b.v);
    for (int k = 0; k < 3; ++k) {
        d.g[k] = scale * (a.g[k] + sign * b.g[k]);
        d.t[k] = scale * (a.t[k] + sign * b.t[k]);
    }
    for (int k = 0; k < 6; ++k)
        d.h[k] = scale * (a.h[k] + sign * b.h[k]);
    d.lap = d.h[0] + d.h[3] + d.h[5];
    d.tau = scale * (a.tau + sign * b.tau);
    return d;
}

// Component k (0 total, 1 alpha, 2 beta, 3 difference) of a point from its
// two channels: (alpha, beta) or (total, difference).
__device__ inline DDv component(const DDv& c0, const DDv& c1, int from_spin, int k) {
    if (from_spin) {
        if (k == 1) return c0;
        if (k == 2) return c1;
        return combine(c0, c1, k == 0 ? 1.0 : -1.0, 1.0);
    }
    if (k == 0) return c0;
    if (k == 3) return c1;
    return combine(c0, c1, k == 1 ? 1.0 : -1.0, 0.5);
}

__device__ inline double gnorm2(const DDv& d) { return (d.g[0] * d.g[0] + d.g[1] * d.g[1]) + d.g[2] * d.g[2]; }


// elf(), lol(), gradient_expansion_locator() of src/evaluator.cpp:127-215
// over NC (1 or 2) channels. Inlined with unrolled loops, so that the channel
// arrays stay in registers.
template <int NC> __device__ __attribute__((always_inline)) inline double elf(const DDv (&c)[NC], double kFermi) {
    double pauli = 0.0, uniform = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!(c[i].v > 0.0)) continue;
        pauli += c[i].tau - gnorm2(c[i]) / (8.0 * c[i].v);
        uniform += kFermi * pow(c[i].v, 5.0 / 3.0);
    }
    if (!(uniform > 0.0)) return 0.0;
    const double chi = pauli / uniform;
    return 1.0 / (1.0 + chi * chi);
}

template <int NC> __device__ __attribute__((always_inline)) inline double lol(const DDv (&c)[NC], double kFermi) {
    double tau = 0.0, uniform = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        tau += c[i].tau;
        if (c[i].v > 0.0) uniform += kFermi * pow(c[i].v, 5.0 / 3.0);
    }
    if (!(uniform > 0.0)) return 0.0;
    if (!(tau > 0.0)) return 1.0;
    return uniform / (tau + uniform);
}

template <int NC>
__device__ __attribute__((always_inline)) inline double ge_locator(const DDv (&c)[NC], bool is_lol, double kFermi) {
    double uniform = 0.0, pauli = 0.0, kinetic = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!(c[i].v > 0.0)) continue;
        const double ref = kFermi * pow(c[i].v, 5.0 / 3.0);
        const double wz = gnorm2(c[i]) / c[i].v;
        uniform += ref;
        pauli += ref - wz / 9.0 + c[i].lap / 6.0;
        kinetic += ref + wz / 72.0 + c[i].lap / 6.0;
    }
    if (!(uniform > 0.0)) return 0.0;
    if (is_lol) {
        const double den = kinetic + uniform;
        return den != 0.0 ? uniform / den : 0.0;
    }
    const double chi = pauli / uniform;
    return 1.0 / (1.0 + chi * chi);
}

// Middle eigenvalue of the symmetric Hessian (xx, xy, xz, yy, yz, zz) by
// cyclic Jacobi rotations in FP64 (the sign decides sign(lambda2) rho). The
// rotation loops are unrolled, so that the matrix stays in registers.
__device__ __attribute__((always_inline)) inline double middle_eigenvalue(const double (&h6)[6]) {
    double a[3][3] = {{h6[0], h6[1], h6[2]}, {h6[1], h6[3], h6[4]}, {h6[2], h6[4], h6[5]}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off == 0.0) break;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
    }
    const double e0 = a[0][0], e1 = a[1][1], e2 = a[2][2];
    if ((e0 <= e1 && e1 <= e2) || (e2 <= e1 && e1 <= e0)) return e1;
    if ((e1 <= e0 && e0 <= e2) || (e2 <= e0 && e0 <= e1)) return e0;
    return e2;
}

__global__ void __launch_bounds__(256) fields_kernel(FieldsLaunch f) {
    const long long j = blockIdx.x * 256ll + threadIdx.x;
    if (j >= f.n) return;
    const long long n = f.n;
    const long long o = f.out_offset + j;
    const DDv c0 = load_channel(f.channels, n, j);
    const DDv c1 = f.closed ? c0 : load_channel(f.channels + kMoments * n, n, j);
    for (int t = 0; t < f.ntasks; ++t) {
        const FieldTask& task = f.tasks[t];
        const int kind = task.kind;
        // Total ELF / LOL / Tsirelson-Stash forms use both spin channels.
        if ((kind == 6 || kind == 10 || kind == 35 || kind == 36) && task.component == 0) {
            const DDv c[2] = {component(c0, c1, f.from_spin, 1), component(c0, c1, f.from_spin, 2)};
            double v;
            if (kind == 6) v = elf(c, f.fermi);
            else if (kind == 10) v = lol(c, f.fermi);
            else v = ge_locator(c, kind == 36, f.fermi);
            task.out[0][o] = v;
            continue;
        }
        if (kind == 16) { // spin polarization
            const double tv = component(c0, c1, f.from_spin, 0).v;
            const double sv = component(c0, c1, f.from_spin, 3).v;
            task.out[0][o] = tv > 0.0 ? sv / tv : 0.0;
            continue;
        }
        const DDv one[1] = {component(c0, c1, f.from_spin, task.component)};
        const DDv& d = one[0];
        double* out0 = task.out[0];
        switch (kind) {
        case 0: out0[o] = d.v; break;
        case 1: for (int a = 0; a < 3; ++a) task.out[a][o] = d.g[a]; break;
        case 2: for (int k = 0; k < 6; ++k) task.out[k][o] = d.h[k]; break;
        case 3: out0[o] = d.lap; break;
        case 4: out0[o] = d.tau; break;
        case 6: out0[o] = elf(one, f.fermi); break;
        case 7: { // reduced gradient
            const double scale = f.rdg_scale;
            double v = 100.0;
            if (d.v > 0.0) {
                const double den = scale * pow(d.v, 4.0 / 3.0);
                const double norm = sqrt(gnorm2(d));
                if (den > 0.0 && norm < 100.0 * den) v = norm / den;
            }
            out0[o] = v;
            break;
        }
        case 8: out0[o] = middle_eigenvalue(d.h) < 0.0 ? -d.v : d.v; break;
        case 10: out0[o] = lol(one, f.fermi); break;
        case 35: out0[o] = ge_locator(one, false, f.fermi); break;
        case 36: out0[o] = ge_locator(one, true, f.fermi); break;
        case 11: { // IRI
            double v = 100.0;
            if (d.v > 0.0) {
                v = sqrt(gnorm2(d)) / pow(d.v, 1.1);
                v = v < 100.0 ? v : 100.0;
            }
            out0[o] = v;
            break;
        }
        case 12: { // DORI
            double v = 0.0;
            if (d.v > 0.0) {
                const double q[3] = {d.g[0] / d.v, d.g[1] / d.v, d.g[2] / d.v};
                const double q2 = (q[0] * q[0] + q[1] * q[1]) + q[2] * q[2];
                if (!(q2 > 0.0)) {
                    v = 1.0;
                } else {
                    const double H[3][3] = {{d.h[0], d.h[1], d.h[2]}, {d.h[1], d.h[3], d.h[4]}, {d.h[2], d.h[4], d.h[5]}};
                    double w[3];
                    for (int a = 0; a < 3; ++a)
                        w[a] = (H[a][0] * q[0] + H[a][1] * q[1] + H[a][2] * q[2]) / d.v - q2 * q[a];
                    const double theta = 4.0 * ((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]) / (q2 * q2 * q2);
                    v = isfinite(theta) ? theta / (1.0 + theta) : 1.0;
                }
            }
            out0[o] = v;
            break;
        }
        case 13: out0[o] = d.tau - 0.25 * d.lap; break;
        case 14: out0[o] = 0.25 * d.lap - 2.0 * d.tau; break;
        case 15: out0[o] = 0.25 * d.lap - d.tau; break;
        case 30: for (int a = 0; a < 3; ++a) task.out[a][o] = d.t[a]; break;
        case 31: {
            const double diag[3] = {d.h[0], d.h[3], d.h[5]};
            for (int a = 0; a < 3; ++a) task.out[a][o] = d.t[a] - 0.25 * diag[a];
            break;
        }
        case 32: out0[o] = d.tau > 0.0 ? fabs(0.25 * d.lap - 2.0 * d.tau) / d.tau : 0.0; break;
        case 33: out0[o] = d.v > 0.0 ? (0.25 * d.lap - d.tau) / d.v : 0.0; break;
        case 34: out0[o] = d.v > 0.0 ? d.tau / d.v : 0.0; break;
        case 17: out0[o] = d.v > 0.0 ? -d.v * log(d.v) : 0.0; break;
        case 18: out0[o] = d.v > 0.0 ? gnorm2(d) / d.v : 0.0; break;
        default: break;
        }
    }
}

__global__ void add_kernel(double* dst, const double* src, long long n) {
    const long long j = blockIdx.x * 256ll + threadIdx.x;
    if (j < n) dst[j] += src[j];
}

__global__ void fill_kernel(double* dst, long long n, double v) {
    const long long j = blockIdx.x * 256ll + threadIdx.x;
    if (j < n) dst[j] = v;
}

// store_orbital_energy_block (src/evaluator.cpp:481-520) per point.
__global__ void orbital_energy_fields_kernel(const double* __restrict__ sums, long long n, long long out_offset,
                                             int ntasks, const FieldTask* __restrict__ tasks, int virtuals,
                                             int fukui) {
    const long long j = blockIdx.x * 256ll + threadIdx.x;
    if (j >= n) return;
    const double rho = sums[kOeOccDensity * n + j];
    const double alie = rho > 0.0 ? sums[kOeOccEnergy * n + j] / rho : 0.0;
    double lea = 0.0, attachment = 0.0, plus = 0.0, minus = 0.0;
    if (virtuals) {
        const double vir = sums[kOeVirDensity * n + j];
        lea = vir > 0.0 ? -sums[kOeVirEnergy * n + j] / vir : 0.0;
        attachment = rho > 0.0 ? sums[kOeAttachment * n + j] / rho : 0.0;
    }
    if (fukui) {
        plus = sums[kOeFukuiPlus * n + j];
        minus = sums[kOeFukuiMinus * n + j];
    }
    for (int t = 0; t < ntasks; ++t) {
        double v;
        switch (tasks[t].kind) {
        case kAlie: v = alie; break;
        case kLea: v = lea; break;
        case kLeae: v = attachment; break;
        case kElectronegativity: v = 0.5 * (alie + lea); break;
        case kHardness: v = 0.5 * (alie - lea); break;
        case kFukuiPlus: v = plus; break;
        case kFukuiMinus: v = minus; break;
        case kFukuiZero: v = 0.5 * (plus + minus); break;
        default: v = plus - minus; break;
        }
        tasks[t].out[0][out_offset + j] = v;
    }
}

// FP64 rescue of ill-conditioned ELF values. The Pauli term tau - |grad
// rho|^2 / (8 rho) of each spin channel is a difference of two positive
// terms; with a relative error eps on the contraction's moments it carries
// an error up to eps (tau + |grad rho|^2 / (8 rho)). A point is recomputed in
// FP64 when that error, propagated through ELF = 1 / (1 + chi^2) (the largest
// |dELF/dchi| = 2 chi / (1 + chi^2)^2 over chi +- dchi), could exceed tol.
__device__ inline bool elf_uncertain(const DDv* c, int nc, double fermi, double eps, double tol) {
    double P = 0.0, Sum = 0.0, U = 0.0;
    bool any = false;
    for (int i = 0; i < nc; ++i) {
        if (!(c[i].v > 0.0)) continue;
        any = true;
        const double q = gnorm2(c[i]) / (8.0 * c[i].v);
        P += c[i].tau - q;
        Sum += c[i].tau + q;
        U += fermi * pow(c[i].v, 5.0 / 3.0);
    }
    if (!any) return false; // ELF is 0 in every precision
    if (!(U > 0.0)) return true;
    const double chi = fabs(P / U), dchi = eps * Sum / U;
    const double lo = fmax(chi - dchi, 0.0), hi = chi + dchi;
    const auto slope = [](double x) { return 2.0 * x / ((1.0 + x * x) * (1.0 + x * x)); };
    const double peak = 0.5773502691896258; // 1/sqrt(3), where the slope is largest
    const double smax = (lo <= peak && peak <= hi) ? slope(peak) : fmax(slope(lo), slope(hi));
    return !(smax * dchi <= tol);
}

// ELF test from each channel's own error scales. With the contraction's
// error model |d psi_i| <= e a_i and |d d_a psi_i| <= e b_ai (e the model's
// unit, calibrated by the caller; a_i = sum_k |C_ik Phi_k|, b_ai the same
// for d_a Phi_k),
// s >= sqrt(sum_i a_i^2) and w >= sum_a sqrt(sum_i b_ai^2) (the triangle
// inequality over the operand rows), Cauchy-Schwarz over the MOs gives
//   |d rho| <= 2 e sqrt(rho) s + (e s)^2,
//   |d tau| <= e sqrt(2 tau) w + (e w)^2 / 2,
//   |d grad rho| <= 2 e (sqrt(rho) w + sqrt(3) sqrt(2 tau) s) + 2 e^2 s w.
// The Pauli term P = tau - |grad rho|^2 / (8 rho) and the uniform-gas term
// U = fermi rho^(5/3) are bounded over those intervals, chi = |P| / U over
// theirs, and, ELF = 1 / (1 + chi^2) being monotone in chi, the point is
// uncertain when the ELF interval is wider than tol on either side.
__device__ inline bool elf_uncertain_scaled(const DDv* c, const float* s, const float* w, int nc, double fermi,
                                            double e, double tol) {
    double P = 0.0, U = 0.0, dP = 0.0, dU = 0.0;
    bool any = false;
    for (int i = 0; i < nc; ++i) {
        const double r = c[i].v;
        if (!(r > 0.0)) continue;
        any = true;
        const double es = e * static_cast<double>(s[i]), ew = e * static_cast<double>(w[i]);
        const double sr = sqrt(r), st = sqrt(2.0 * fmax(c[i].tau, 0.0));
        const double drho = 2.0 * sr * es + es * es;
        if (!(drho < r)) return true;
        const double dtau = st * ew + 0.5 * ew * ew;
        const double dg = 2.0 * (sr * ew + 1.7320508075688772 * st * es) + 2.0 * es * ew;
        const double g2 = gnorm2(c[i]), gn = sqrt(g2), q = g2 / (8.0 * r);
        const double ghi = gn + dg, glo = fmax(gn - dg, 0.0);
        const double dq = fmax(ghi * ghi / (8.0 * (r - drho)) - q, q - glo * glo / (8.0 * (r + drho)));
        P += c[i].tau - q;
        dP += dtau + dq;
        const double u = fermi * pow(r, 5.0 / 3.0);
        U += u;
        dU += fermi * pow(r + drho, 5.0 / 3.0) - u;
    }
    if (!any) return false; // ELF is 0 in every precision
    if (!(U - dU > 0.0)) return true;
    const double chi = fabs(P) / U;
    const double lo = fmax(fabs(P) - dP, 0.0) / (U + dU), hi = (fabs(P) + dP) / (U - dU);
    const double elf = 1.0 / (1.0 + chi * chi);
    const double widest = fmax(1.0 / (1.0 + lo * lo) - elf, elf - 1.0 / (1.0 + hi * hi));
    return !(widest <= tol);
}

// Relative density error bound of one channel, 2 eps s / sqrt(rho)
// (0 where the channel holds no density).
__device__ inline double rho_rel_bound(double rho, double s, double eps) {
    return rho > 0.0 ? 2.0 * eps * s / sqrt(rho) : 0.0;
}

// Output bits: 0 total (both channels), 1 alpha, 2 beta.
__global__ void rescue_flag_kernel(const double* ch, long long n, int closed, RescueCriteria rc, double fermi,
                                   uint8_t* flag) {
    const long long j = blockIdx.x * 256ll + threadIdx.x;
    if (j >= n) return;
    DDv c[2];
    c[0] = load_channel(ch, n, j);
    c[1] = closed ? c[0] : load_channel(ch + kMoments * n, n, j);
    bool f = false;
    const int ew = rc.elf_which;
    if (rc.grad_scale) {
        const float sv[2] = {rc.rho_scale[j], closed ? rc.rho_scale[j] : rc.rho_scale[n + j]};
        const float gv[2] = {rc.grad_scale[j], closed ? rc.grad_scale[j] : rc.grad_scale[n + j]};
        const double e = rc.elf_scale_eps;
        if (ew & 1) f = f || elf_uncertain_scaled(c, sv, gv, 2, fermi, e, rc.elf_tol);
        if (ew & 2) f = f || elf_uncertain_scaled(c, sv, gv, 1, fermi, e, rc.elf_tol);
        if (ew & 4) f = f || elf_uncertain_scaled(c + 1, sv + 1, gv + 1, 1, fermi, e, rc.elf_tol);
    } else {
        if (ew & 1) f = f || elf_uncertain(c, 2, fermi, rc.elf_eps, rc.elf_tol);
        if (ew & 2) f = f || elf_uncertain(c, 1, fermi, rc.elf_eps, rc.elf_tol);
        if (ew & 4) f = f || elf_uncertain(c + 1, 1, fermi, rc.elf_eps, rc.elf_tol);
    }
    if (rc.rho_which) {
        const double sa = rc.rho_scale[j], sb = closed ? sa : rc.rho_scale[n + j];
        const double ba = rho_rel_bound(c[0].v, sa, rc.rho_eps), bb = rho_rel_bound(c[1].v, sb, rc.rho_eps);
        const double rt = c[0].v + c[1].v;
        const double bt = rt > 0.0 ? (c[0].v * ba + c[1].v * bb) / rt : 0.0;
        f = f || ((rc.rho_which & 1) && bt > rc.rho_tol) || ((rc.rho_which & 2) && ba > rc.rho_tol) ||
            ((rc.rho_which & 4) && bb > rc.rho_tol);
    }
    flag[j] = f ? 1 : 0;
}

// Per screening block (one workgroup each): the flagged points, compacted in
// ascending order into offs[b * block_size ...], and their count.
__global__ void __launch_bounds__(256) rescue_compact_kernel(const uint8_t* flag, long long n, int block_size,
                                                             uint16_t* offs, int32_t* count) {
    __shared__ int wsum[4];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long long b0 = static_cast<long long>(blockIdx.x) * block_size;
    const int bn = static_cast<int>(min(static_cast<long long>(block_size), n - b0));
    if (tid == 0) base = 0;
    __syncthreads();
    for (int i0 = 0; i0 < bn; i0 += 256) {
        const int i = i0 + tid;
        const bool f = i < bn && flag[b0 + i];
        const unsigned long long ballot = __ballot(f);
        const int before = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(ballot);
        __syncthreads();
        int off = base;
        for (int k = 0; k < w; ++k)
            off += wsum[k];
        if (f) offs[b0 + off + before] = static_cast<uint16_t>(i);
        __syncthreads();
        if (tid == 0) base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (tid == 0) count[blockIdx.x] = base;
}

// Tiles of `tile` flagged points: tile t covers points [16 sub, ...) of the
// flagged list of block tile_blk[t]; total[0] receives the number of tiles,
// total[1] the number of flagged points.
__global__ void __launch_bounds__(1024) rescue_tiles_kernel(const int32_t* count, int nblk, int tile,
                                                            int32_t* tile_blk, int32_t* tile_sub, int32_t* total) {
    // Each thread takes a run of consecutive blocks; one scan over the
    // workgroup of the runs' tile counts gives every run its first tile.
    __shared__ int wsum[16], fsum[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int per = (nblk + 1023) / 1024;
    const int b0 = min(nblk, tid * per), b1 = min(nblk, b0 + per);
    int local = 0, flagged = 0;
    for (int b = b0; b < b1; ++b) {
        local += (count[b] + tile - 1) / tile;
        flagged += count[b];
    }
    for (int o = 32; o > 0; o >>= 1)
        flagged += __shfl_xor(flagged, o, 64);
    if (lane == 0) fsum[w] = flagged;
    int x = local;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int off = x - local;
    for (int k = 0; k < w; ++k)
        off += wsum[k];
    for (int b = b0; b < b1; ++b) {
        const int nt = (count[b] + tile - 1) / tile;
        for (int s2 = 0; s2 < nt; ++s2) {
            tile_blk[off + s2] = b;
            tile_sub[off + s2] = s2;
        }
        off += nt;
    }
    if (tid == 1023) total[0] = off; // the last run ends at the total
    if (tid == 0) {
        int f = 0;
        for (int k = 0; k < 16; ++k)
            f += fsum[k];
        total[1] = f;
    }
}

// Writes the rescued channel moments (summed over MO slot groups) over the
// contraction's values at the flagged points.
__global__ void rescue_scatter_kernel(const double* rm, long long rslots, int groups, int tile, int block_size,
                                      const int32_t* tile_blk, const int32_t* tile_sub, const int32_t* total,
                                      const int32_t* count, const uint16_t* offs, double* ch, long long n) {
    const long long q = blockIdx.x * 256ll + threadIdx.x;
    const long long t = q / tile;
    const int jj = static_cast<int>(q % tile);
    if (t >= *total) return;
    const int b = tile_blk[t], sub = tile_sub[t];
    if (sub * tile + jj >= count[b]) return;
    const long long point = static_cast<long long>(b) * block_size + offs[static_cast<long long>(b) * block_size + sub * tile + jj];
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < kMoments; ++k) {
            double v = 0.0;
            for (int g = 0; g < groups; ++g)
                v += rm[((static_cast<long long>(g) * 2 + c) * kMoments + k) * rslots + q];
            ch[(static_cast<long long>(c) * kMoments + k) * n + point] = v;
        }
}

} // namespace

void launch_rescue_plan(const double* channel_moments, long long n, int closed, const RescueCriteria& c,
                        int block_size, int nblk, int tile, const RescuePlan& r, hipStream_t stream) {
    if (n <= 0) return;
    if (c.rho_which && !c.rho_scale) invalid("density rescue without error scales");
    const FieldsLaunch f = fields_launch(nullptr, 0, 0, 0, 0, 0, nullptr);
    rescue_flag_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(channel_moments, n, closed, c,
                                                                                  f.fermi, r.flag);
    KUMO_HIP(hipGetLastError());
    rescue_compact_kernel<<<static_cast<unsigned>(nblk), 256, 0, stream>>>(r.flag, n, block_size, r.offs, r.count);
    KUMO_HIP(hipGetLastError());
    rescue_tiles_kernel<<<1, 1024, 0, stream>>>(r.count, nblk, tile, r.tile_blk, r.tile_sub, r.total);
    KUMO_HIP(hipGetLastError());
}

void launch_rescue_scatter(const double* rescue_moments, long long rslots, int groups, int tile, int block_size,
                           const RescuePlan& r, double* channel_moments, long long n, hipStream_t stream) {
    if (rslots <= 0) return;
    rescue_scatter_kernel<<<static_cast<unsigned>((rslots + 255) / 256), 256, 0, stream>>>(
        rescue_moments, rslots, groups, tile, block_size, r.tile_blk, r.tile_sub, r.total, r.count, r.offs,
        channel_moments, n);
    KUMO_HIP(hipGetLastError());
}

// Field tasks passed by value (kernel arguments are part of the launch, so a
// caller capturing a graph needs no host buffer to outlive the call).
struct TaskPack {
    FieldTask t[kMaxStoredTasks];
};

__global__ void store_tasks_kernel(TaskPack pack, int n, FieldTask* dst) {
    const int i = threadIdx.x;
    if (i < n) dst[i] = pack.t[i];
}

void launch_store_tasks(const FieldTask* tasks, int n, FieldTask* dst, hipStream_t stream) {
    if (n <= 0) return;
    if (n > kMaxStoredTasks) invalid("at most " + std::to_string(kMaxStoredTasks) + " point-path properties per call");
    TaskPack pack{};
    for (int i = 0; i < n; ++i)
        pack.t[i] = tasks[i];
    store_tasks_kernel<<<1, 64, 0, stream>>>(pack, n, dst);
    KUMO_HIP(hipGetLastError());
}

void launch_add_into(double* dst, const double* src, long long n, hipStream_t stream) {
    if (n <= 0) return;
    add_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(dst, src, n);
    KUMO_HIP(hipGetLastError());
}

void launch_fill(double* dst, long long n, double value, hipStream_t stream) {
    if (n <= 0) return;
    fill_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(dst, n, value);
    KUMO_HIP(hipGetLastError());
}

void launch_orbital_energy_fields(const double* sums, long long n, long long out_offset, int ntasks,
                                  const FieldTask* tasks, int virtuals, int fukui, hipStream_t stream) {
    if (n <= 0 || ntasks == 0) return;
    orbital_energy_fields_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(
        sums, n, out_offset, ntasks, tasks, virtuals, fukui);
    KUMO_HIP(hipGetLastError());
}

void launch_fields(const FieldsLaunch& f, hipStream_t stream) {
    if (f.n <= 0 || f.ntasks == 0) return;
    fields_kernel<<<static_cast<unsigned>((f.n + 255) / 256), 256, 0, stream>>>(f);
    KUMO_HIP(hipGetLastError());
}

} // namespace kumo
