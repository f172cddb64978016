// Device evaluation of contracted Gaussian AO blocks at one point.
//
// The arithmetic follows the reference term by term
// (src/basis_evaluator.cpp:442-525 for values, :512-643 for derivatives):
// the same products in the same association and no fused multiply-add
// (`#pragma clang fp contract(off)`), so that FP64 AO values equal the CPU
// ones up to the last-bit differences of exp(), and the screening masks built
// from them are the reference's. The shell's angular momentum is a template
// parameter so every per-component array lives in registers.
#pragma once

#include "device.hpp"

#include <hip/hip_runtime.h>

namespace kumo {

struct MolView {
    const DevShell* shells;
    int nshells;
    int nao;
    const double* exps;
    const double* weights;
    const double* tdense; // [kMaxL+1][kMaxPure][kMaxCart] Cartesian->pure, zero padded
    const double* radw;   // per primitive: weight of the block's first Cartesian component
    const double* tg;     // [kMaxL+1][kMaxPure][kMaxCart] T[r][k] g[k] (transform of unscaled components)
    const double* gk;     // [kMaxL+1][kMaxCart] g[k]: Cartesian AO = g[k] x unscaled component
    const int32_t* lorder; // blocks by descending l
    const ShellRec* rec;   // per block, for the contraction kernels
};

// Cartesian powers in Gaussian FCHK order (src/basis_evaluator.cpp:135-152).
__device__ constexpr signed char kPow[kMaxL + 1][kMaxCart][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}},
    {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0}, {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}},
    {{0, 0, 4}, {0, 1, 3}, {0, 2, 2}, {0, 3, 1}, {0, 4, 0}, {1, 0, 3}, {1, 1, 2}, {1, 2, 1}, {1, 3, 0}, {2, 0, 2},
     {2, 1, 1}, {2, 2, 0}, {3, 0, 1}, {3, 1, 0}, {4, 0, 0}},
    {{0, 0, 5}, {0, 1, 4}, {0, 2, 3}, {0, 3, 2}, {0, 4, 1}, {0, 5, 0}, {1, 0, 4}, {1, 1, 3}, {1, 2, 2}, {1, 3, 1},
     {1, 4, 0}, {2, 0, 3}, {2, 1, 2}, {2, 2, 1}, {2, 3, 0}, {3, 0, 2}, {3, 1, 1}, {3, 2, 0}, {4, 0, 1}, {4, 1, 0},
     {5, 0, 0}}};

// Panel sets: bit c selects derivative component c of
// value, x, y, z, xx, xy, xz, yy, yz, zz (src/derivative_layout.hpp:13), and
// bit 10 the Laplacian xx + yy + zz (contractions that need only the trace).
constexpr unsigned kPanelsValue = 0x001;
constexpr unsigned kPanelsGrad = 0x00f;
constexpr unsigned kPanelsDiag = 0x08f | 0x010 | 0x200; // value, grad, xx, yy, zz
constexpr unsigned kPanelsAll = 0x3ff;
constexpr unsigned kPanelsLap = 0x40f;                   // value, grad, Laplacian
constexpr int kPanelLap = 10;

__host__ __device__ constexpr int popcount10(unsigned m) {
    int n = 0;
    for (int i = 0; i <= kPanelLap; ++i)
        n += (m >> i) & 1u;
    return n;
}
// Index of the derivative component of output panel i.
__host__ __device__ constexpr int panel_component(unsigned m, int i) {
    for (int c = 0; c <= kPanelLap; ++c)
        if ((m >> c) & 1u) {
            if (i == 0) return c;
            --i;
        }
    return -1;
}

// Grid or listed point q (src/grid_geometry.hpp:59-68): origin + steps*(i,j,k),
// each coordinate summed as ((s0 i + s1 j) + s2 k) without fused operations.
__device__ __attribute__((always_inline)) inline void point_coords(const PointSource& ps, long long q, double& x, double& y, double& z) {
#pragma clang fp contract(off)
    if (ps.points) {
        x = ps.points[3 * q];
        y = ps.points[3 * q + 1];
        z = ps.points[3 * q + 2];
        return;
    }
    const long long plane = ps.ny * ps.nz;
    const double fi = static_cast<double>(q / plane);
    const double fj = static_cast<double>((q / ps.nz) % ps.ny);
    const double fk = static_cast<double>(q % ps.nz);
    x = ps.origin[0] + ((ps.steps[0] * fi + ps.steps[3] * fj) + ps.steps[6] * fk);
    y = ps.origin[1] + ((ps.steps[1] * fi + ps.steps[4] * fj) + ps.steps[7] * fk);
    z = ps.origin[2] + ((ps.steps[2] * fi + ps.steps[5] * fj) + ps.steps[8] * fk);
}

// point_coords of point q0 + d (0 <= d < 2^30): the grid indices of q0 are
// found once (the 64-bit divisions are the same for every lane of a tile when
// q0 is uniform) and advanced by d in 32-bit arithmetic; the coordinates are
// point_coords' bit for bit.
__device__ __attribute__((always_inline)) inline void point_coords_from(const PointSource& ps, long long q0, int d,
                                                                      double& x, double& y, double& z) {
#pragma clang fp contract(off)
    if (ps.points) {
        point_coords(ps, q0 + d, x, y, z);
        return;
    }
    const long long plane = ps.ny * ps.nz;
    const long long i0 = q0 / plane, r0 = q0 - i0 * plane;
    const long long j0 = r0 / ps.nz, k0 = r0 - j0 * ps.nz;
    // k0 + d < nz + 2^30 and j0 + dj < ny + 2^30 fit 32 bits when nz, ny < 2^31
    if (ps.nz >= (1ll << 31) || ps.ny >= (1ll << 31)) {
        point_coords(ps, q0 + d, x, y, z);
        return;
    }
    const unsigned nz = static_cast<unsigned>(ps.nz), ny = static_cast<unsigned>(ps.ny);
    const unsigned t = static_cast<unsigned>(k0) + static_cast<unsigned>(d);
    const unsigned dj = t / nz, k = t - dj * nz;
    const unsigned tj = static_cast<unsigned>(j0) + dj;
    const unsigned di = tj / ny, jj = tj - di * ny;
    const double fi = static_cast<double>(i0 + di);
    const double fj = static_cast<double>(jj);
    const double fk = static_cast<double>(k);
    x = ps.origin[0] + ((ps.steps[0] * fi + ps.steps[3] * fj) + ps.steps[6] * fk);
    y = ps.origin[1] + ((ps.steps[1] * fi + ps.steps[4] * fj) + ps.steps[7] * fk);
    z = ps.origin[2] + ((ps.steps[2] * fi + ps.steps[5] * fj) + ps.steps[8] * fk);
}

__device__ __attribute__((always_inline)) inline double sq_norm(double x, double y, double z) {
#pragma clang fp contract(off)
    return (x * x + y * y) + z * z;
}

// Cartesian components of block `s` (angular momentum L) at displacement
// (x, y, z), r2 = |r|^2: cart[k][i] for output panel i of PSET.
// ORDER 0 uses the value kernel (evaluate_into_impl), 1 and 2 the derivative
// kernel (evaluate_derivatives_batch), as the reference does.
template <int L, int ORDER, unsigned PSET>
__device__ __attribute__((always_inline)) inline void eval_cartesian(const DevShell& s, const double* __restrict__ exps,
                                      const double* __restrict__ wts, double x, double y, double z, double r2,
                                      double (&cart)[(L + 1) * (L + 2) / 2][popcount10(PSET)]) {
#pragma clang fp contract(off)
    static_assert(!(PSET >> kPanelLap), "reference-order evaluation has no Laplacian panel");
    constexpr int NC = (L + 1) * (L + 2) / 2;
    constexpr int NP = popcount10(PSET);
    const int np = s.nprim;
    const double* e = exps + s.prim;
    const double* w = wts + s.wofs;
    double xp[L + 1], yp[L + 1], zp[L + 1];
    xp[0] = yp[0] = zp[0] = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) {
        xp[n] = xp[n - 1] * x;
        yp[n] = yp[n - 1] * y;
        zp[n] = zp[n - 1] * z;
    }
    if constexpr (ORDER == 0) {
        double rs[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k)
            rs[k] = 0.0;
        for (int p = 0; p < np; ++p) {
            const double radial = exp(-e[p] * r2);
#pragma unroll
            for (int k = 0; k < NC; ++k)
                rs[k] += w[k * np + p] * radial;
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const double poly = (xp[kPow[L][k][0]] * yp[kPow[L][k][1]]) * zp[kPow[L][k][2]];
            cart[k][0] = poly * rs[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int i = 0; i < NP; ++i)
                cart[k][i] = 0.0;
        const double r[3] = {x, y, z};
        for (int p = 0; p < np; ++p) {
            const double a = e[p];
            const double radial = exp(-a * r2);
            // f[axis][m][q]: x^m exp(-a x^2) and its derivatives over exp(-a x^2).
            double f[3][L + 1][3];
#pragma unroll
            for (int axis = 0; axis < 3; ++axis) {
                const double* pw = axis == 0 ? xp : (axis == 1 ? yp : zp);
                const double c = r[axis];
#pragma unroll
                for (int m = 0; m <= L; ++m) {
                    const double xm = pw[m];
                    f[axis][m][0] = xm;
                    f[axis][m][1] = (m ? m * pw[m > 0 ? m - 1 : 0] : 0.0) - 2.0 * a * c * xm;
                    if constexpr (ORDER == 2)
                        f[axis][m][2] = (m >= 2 ? m * (m - 1) * pw[m >= 2 ? m - 2 : 0] : 0.0) -
                                        2.0 * a * (2 * m + 1) * xm + 4.0 * a * a * c * c * xm;
                    else
                        f[axis][m][2] = 0.0;
                }
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const double wgt = w[k * np + p] * radial;
                const double* fx = f[0][kPow[L][k][0]];
                const double* fy = f[1][kPow[L][k][1]];
                const double* fz = f[2][kPow[L][k][2]];
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    constexpr int dummy = 0;
                    (void)dummy;
                    const int c = panel_component(PSET, i);
                    double t;
                    switch (c) {
                    case 0: t = wgt * fx[0] * fy[0] * fz[0]; break;
                    case 1: t = wgt * fx[1] * fy[0] * fz[0]; break;
                    case 2: t = wgt * fx[0] * fy[1] * fz[0]; break;
                    case 3: t = wgt * fx[0] * fy[0] * fz[1]; break;
                    case 4: t = wgt * fx[2] * fy[0] * fz[0]; break;
                    case 5: t = wgt * fx[1] * fy[1] * fz[0]; break;
                    case 6: t = wgt * fx[1] * fy[0] * fz[1]; break;
                    case 7: t = wgt * fx[0] * fy[2] * fz[0]; break;
                    case 8: t = wgt * fx[0] * fy[1] * fz[1]; break;
                    default: t = wgt * fx[0] * fy[0] * fz[2]; break;
                    }
                    cart[k][i] += t;
                }
            }
        }
    }
}

// AO values of block `s` at the point: out(row, i) for each AO row of the
// block and output panel i; rows are pure (transformed, in increasing
// Cartesian-component order) or Cartesian. Calls emit(row, vals[NP]).
template <int L, int ORDER, unsigned PSET, typename Emit>
__device__ __attribute__((always_inline)) inline void eval_block_l(const DevShell& s, const MolView& m, double x, double y, double z, double r2,
                                    Emit&& emit) {
#pragma clang fp contract(off)
    constexpr int NC = (L + 1) * (L + 2) / 2;
    constexpr int NP = popcount10(PSET);
    double cart[NC][NP];
    eval_cartesian<L, ORDER, PSET>(s, m.exps, m.weights, x, y, z, r2, cart);
    if (L >= 2 && s.pure) {
        const double* T = m.tdense + L * kMaxPure * kMaxCart;
#pragma unroll
        for (int row = 0; row < 2 * L + 1; ++row) {
            double v[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i)
                v[i] = 0.0;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const double t = T[row * kMaxCart + k];
#pragma unroll
                for (int i = 0; i < NP; ++i)
                    v[i] += t * cart[k][i];
            }
            emit(row, v);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NC; ++k)
            emit(k, cart[k]);
    }
}

// Runtime dispatch on the block's angular momentum.
// eval_block_l with the angular momentum a runtime value: the same
// operations in the same order, with the per-component arrays in private
// memory. Used for l >= 4, where unrolled register arrays would set the
// register budget of every kernel that can meet such a block.
template <int ORDER, unsigned PSET, typename Emit>
__device__ __attribute__((always_inline)) inline void eval_block_rt(const DevShell& s, const MolView& m, double x, double y,
                                                                  double z, double r2, Emit&& emit) {
#pragma clang fp contract(off)
    static_assert(!(PSET >> kPanelLap), "reference-order evaluation has no Laplacian panel");
    constexpr int NP = popcount10(PSET);
    const int L = s.l, NC = s.ncart, np = s.nprim;
    const double* e = m.exps + s.prim;
    const double* w = m.weights + s.wofs;
    double cart[kMaxCart][NP];
    double xp[kMaxL + 1], yp[kMaxL + 1], zp[kMaxL + 1];
    xp[0] = yp[0] = zp[0] = 1.0;
    for (int n = 1; n <= L; ++n) {
        xp[n] = xp[n - 1] * x;
        yp[n] = yp[n - 1] * y;
        zp[n] = zp[n - 1] * z;
    }
    if constexpr (ORDER == 0) {
        double rs[kMaxCart];
        for (int k = 0; k < NC; ++k)
            rs[k] = 0.0;
        for (int p = 0; p < np; ++p) {
            const double radial = exp(-e[p] * r2);
            for (int k = 0; k < NC; ++k)
                rs[k] += w[k * np + p] * radial;
        }
        for (int k = 0; k < NC; ++k) {
            const double poly = (xp[kPow[L][k][0]] * yp[kPow[L][k][1]]) * zp[kPow[L][k][2]];
            cart[k][0] = poly * rs[k];
        }
    } else {
        for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int i = 0; i < NP; ++i)
                cart[k][i] = 0.0;
        const double r[3] = {x, y, z};
        for (int p = 0; p < np; ++p) {
            const double a = e[p];
            const double radial = exp(-a * r2);
            double f[3][kMaxL + 1][3];
            for (int axis = 0; axis < 3; ++axis) {
                const double* pw = axis == 0 ? xp : (axis == 1 ? yp : zp);
                const double c = r[axis];
                for (int mm = 0; mm <= L; ++mm) {
                    const double xm = pw[mm];
                    f[axis][mm][0] = xm;
                    f[axis][mm][1] = (mm ? mm * pw[mm > 0 ? mm - 1 : 0] : 0.0) - 2.0 * a * c * xm;
                    if constexpr (ORDER == 2)
                        f[axis][mm][2] = (mm >= 2 ? mm * (mm - 1) * pw[mm >= 2 ? mm - 2 : 0] : 0.0) -
                                         2.0 * a * (2 * mm + 1) * xm + 4.0 * a * a * c * c * xm;
                    else
                        f[axis][mm][2] = 0.0;
                }
            }
            for (int k = 0; k < NC; ++k) {
                const double wgt = w[k * np + p] * radial;
                const double* fx = f[0][kPow[L][k][0]];
                const double* fy = f[1][kPow[L][k][1]];
                const double* fz = f[2][kPow[L][k][2]];
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const int c = panel_component(PSET, i);
                    double t;
                    switch (c) {
                    case 0: t = wgt * fx[0] * fy[0] * fz[0]; break;
                    case 1: t = wgt * fx[1] * fy[0] * fz[0]; break;
                    case 2: t = wgt * fx[0] * fy[1] * fz[0]; break;
                    case 3: t = wgt * fx[0] * fy[0] * fz[1]; break;
                    case 4: t = wgt * fx[2] * fy[0] * fz[0]; break;
                    case 5: t = wgt * fx[1] * fy[1] * fz[0]; break;
                    case 6: t = wgt * fx[1] * fy[0] * fz[1]; break;
                    case 7: t = wgt * fx[0] * fy[2] * fz[0]; break;
                    case 8: t = wgt * fx[0] * fy[1] * fz[1]; break;
                    default: t = wgt * fx[0] * fy[0] * fz[2]; break;
                    }
                    cart[k][i] += t;
                }
            }
        }
    }
    if (L >= 2 && s.pure) {
        const double* T = m.tdense + L * kMaxPure * kMaxCart;
        for (int row = 0; row < 2 * L + 1; ++row) {
            double v[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i)
                v[i] = 0.0;
            for (int k = 0; k < NC; ++k) {
                const double t = T[row * kMaxCart + k];
#pragma unroll
                for (int i = 0; i < NP; ++i)
                    v[i] += t * cart[k][i];
            }
            emit(row, v);
        }
    } else {
        for (int k = 0; k < NC; ++k) {
            double v[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i)
                v[i] = cart[k][i];
            emit(k, v);
        }
    }
}

template <int ORDER, unsigned PSET, typename Emit>
__device__ __attribute__((always_inline)) inline void eval_block(const DevShell& s, const MolView& m, double x, double y, double z, double r2,
                                  Emit&& emit) {
    switch (s.l) {
    case 0: eval_block_l<0, ORDER, PSET>(s, m, x, y, z, r2, emit); break;
    case 1: eval_block_l<1, ORDER, PSET>(s, m, x, y, z, r2, emit); break;
    case 2: eval_block_l<2, ORDER, PSET>(s, m, x, y, z, r2, emit); break;
    case 3: eval_block_l<3, ORDER, PSET>(s, m, x, y, z, r2, emit); break;
    default: eval_block_rt<ORDER, PSET>(s, m, x, y, z, r2, emit); break;
    }
}

// Factored evaluation for the contraction kernels. The contraction weights of
// a block factor as weights[k][p] = g[k] radw[p] (g depends on l and the
// component only), so component k is g[k] P_k(x, y, z) R(r^2) with
// P_k = x^i y^j z^k and R = sum_p radw[p] exp(-a_p r^2). The kernels fold g[k]
// into the MO coefficients and generate the unscaled P_k R and its
// derivatives from three radial sums: R0 = R, R1 = sum_p -2 a_p radw[p] e_p,
// R2 = sum_p 4 a_p^2 radw[p] e_p, with d/dx R0 = x R1 and d/dx R1 = x R2.
template <int ORDER>
__device__ __attribute__((always_inline)) inline void radial_sums(const DevShell& s, const MolView& m, double r2,
                                                                double (&R)[3]) {
#pragma clang fp contract(fast)
    const double* e = m.exps + s.prim;
    const double* u = m.radw + s.prim;
    double r0 = 0.0, r1 = 0.0, rr = 0.0;
    for (int p = 0; p < s.nprim; ++p) {
        const double a = e[p];
        const double g = u[p] * exp(-a * r2);
        r0 += g;
        if constexpr (ORDER >= 1) {
            const double t = -2.0 * a * g;
            r1 += t;
            if constexpr (ORDER == 2) rr += -2.0 * a * t;
        }
    }
    R[0] = r0;
    R[1] = r1;
    R[2] = rr;
}

// radial_sums from a block's record (primitives inline up to kRecPrim).
template <int ORDER>
__device__ __attribute__((always_inline)) inline void radial_sums_rec(const ShellRec& s, const MolView& m, double r2,
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
    const int np = s.nprim;
    if (np <= kRecPrim) {
        for (int p = 0; p < np; ++p)
            add(s.e[p], s.w[p]);
    } else {
        for (int p = 0; p < np; ++p)
            add(m.exps[s.prim + p], m.radw[s.prim + p]);
    }
    R[0] = r0;
    R[1] = r1;
    R[2] = rr;
}

// Panels of the unscaled component P_k R for each k of angular momentum L:
// emit(k, v[NP]). With f0 = c^q, f1 = q c^(q-1), f2 = q (q-1) c^(q-2) per axis:
//   d/da (P R)      = P_a R0 + a P R1
//   d2/da2 (P R)    = P_aa R0 + (2 a P_a + P) R1 + a^2 P R2
//   d2/da db (P R)  = P_ab R0 + (b P_a + a P_b) R1 + a b P R2.
template <int L, int ORDER, unsigned PSET, typename T = double, typename Emit>
__device__ __attribute__((always_inline)) inline void cart_components(T x, T y, T z, const T (&R)[3], Emit&& emit) {
#pragma clang fp contract(fast)
    constexpr int NC = (L + 1) * (L + 2) / 2;
    constexpr int NP = popcount10(PSET);
    const T c[3] = {x, y, z};
    T pw[3][L + 1];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        pw[a][0] = T(1);
#pragma unroll
        for (int n = 1; n <= L; ++n)
            pw[a][n] = pw[a][n - 1] * c[a];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        T f[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int q = kPow[L][k][a];
            f[a][0] = pw[a][q];
            f[a][1] = q >= 1 ? T(q) * pw[a][q >= 1 ? q - 1 : 0] : T(0);
            f[a][2] = q >= 2 ? T(q * (q - 1)) * pw[a][q >= 2 ? q - 2 : 0] : T(0);
        }
        const T P = f[0][0] * f[1][0] * f[2][0];
        T v[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int cc = panel_component(PSET, i);
            if (cc == 0) {
                v[i] = P * R[0];
            } else if (cc <= 3) {
                const int a = cc - 1;
                const T Pa = (a == 0 ? f[0][1] : f[0][0]) * (a == 1 ? f[1][1] : f[1][0]) * (a == 2 ? f[2][1] : f[2][0]);
                v[i] = Pa * R[0] + c[a] * P * R[1];
            } else if (cc == 4 || cc == 7 || cc == 9) {
                const int a = cc == 4 ? 0 : (cc == 7 ? 1 : 2);
                const T Pa = (a == 0 ? f[0][1] : f[0][0]) * (a == 1 ? f[1][1] : f[1][0]) * (a == 2 ? f[2][1] : f[2][0]);
                const T Paa = (a == 0 ? f[0][2] : f[0][0]) * (a == 1 ? f[1][2] : f[1][0]) * (a == 2 ? f[2][2] : f[2][0]);
                v[i] = Paa * R[0] + (T(2) * c[a] * Pa + P) * R[1] + c[a] * c[a] * P * R[2];
            } else if (cc == kPanelLap) {
                // sum of the three above; sum_a c_a P_a = L P for a monomial of degree L
                const T Plap = f[0][2] * f[1][0] * f[2][0] + f[0][0] * f[1][2] * f[2][0] + f[0][0] * f[1][0] * f[2][2];
                v[i] = Plap * R[0] + T(2 * L + 3) * P * R[1] + (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) * P * R[2];
            } else {
                const int a = cc == 8 ? 1 : 0, b = cc == 5 ? 1 : 2;
                const T Pa = (a == 0 ? f[0][1] : f[0][0]) * (a == 1 ? f[1][1] : f[1][0]) * f[2][0];
                const T Pb = f[0][0] * (b == 1 ? f[1][1] : f[1][0]) * (b == 2 ? f[2][1] : f[2][0]);
                const T Pab = (a == 0 ? f[0][1] : f[0][0]) * ((a == 1 || b == 1) ? f[1][1] : f[1][0]) *
                              (b == 2 ? f[2][1] : f[2][0]);
                v[i] = Pab * R[0] + (c[b] * Pa + c[a] * Pb) * R[1] + c[a] * c[b] * P * R[2];
            }
        }
        emit(k, v);
    }
}

// cart_components with the angular momentum and the component index as
// runtime values (a loop over components instead of an unrolled body): for
// the rare high-l blocks and partially kept pure blocks, where unrolling every
// component would cost registers in every other path of a kernel.
template <int ORDER, unsigned PSET, typename T = double>
__device__ __attribute__((always_inline)) inline void cart_component_pw(const int (&qq)[3], T x, T y, T z,
                                                                      const T (&R)[3], T (&v)[popcount10(PSET)]);

template <int ORDER, unsigned PSET, typename T = double>
__device__ __attribute__((always_inline)) inline void cart_component_rt(int L, int k, T x, T y, T z, const T (&R)[3],
                                                                      T (&v)[popcount10(PSET)]) {
    const int qq[3] = {kPow[L][k][0], kPow[L][k][1], kPow[L][k][2]};
    cart_component_pw<ORDER, PSET, T>(qq, x, y, z, R, v);
}

// The component with powers (qq[0], qq[1], qq[2]) as runtime values.
template <int ORDER, unsigned PSET, typename T>
__device__ __attribute__((always_inline)) inline void cart_component_pw(const int (&qq)[3], T x, T y, T z,
                                                                      const T (&R)[3], T (&v)[popcount10(PSET)]) {
#pragma clang fp contract(fast)
    constexpr int NP = popcount10(PSET);
    const T c[3] = {x, y, z};
    T f[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int q = qq[a];
        T p2 = T(1); // c^(q-2)
        for (int n = 2; n < q; ++n)
            p2 *= c[a];
        const T p1 = q >= 2 ? p2 * c[a] : T(1); // c^(q-1)
        f[a][0] = q >= 1 ? p1 * c[a] : T(1);
        f[a][1] = q >= 1 ? T(q) * p1 : T(0);
        f[a][2] = q >= 2 ? T(q * (q - 1)) * p2 : T(0);
    }
    const T P = f[0][0] * f[1][0] * f[2][0];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int cc = panel_component(PSET, i);
        if (cc == 0) {
            v[i] = P * R[0];
        } else if (cc <= 3) {
            const int a = cc - 1;
            const T Pa = (a == 0 ? f[0][1] : f[0][0]) * (a == 1 ? f[1][1] : f[1][0]) * (a == 2 ? f[2][1] : f[2][0]);
            v[i] = Pa * R[0] + c[a] * P * R[1];
        } else if (cc == 4 || cc == 7 || cc == 9) {
            const int a = cc == 4 ? 0 : (cc == 7 ? 1 : 2);
            const T Pa = (a == 0 ? f[0][1] : f[0][0]) * (a == 1 ? f[1][1] : f[1][0]) * (a == 2 ? f[2][1] : f[2][0]);
            const T Paa = (a == 0 ? f[0][2] : f[0][0]) * (a == 1 ? f[1][2] : f[1][0]) * (a == 2 ? f[2][2] : f[2][0]);
            v[i] = Paa * R[0] + (T(2) * c[a] * Pa + P) * R[1] + c[a] * c[a] * P * R[2];
        } else if (cc == kPanelLap) {
            // sum of the three above; sum_a c_a P_a = (qx + qy + qz) P
            const T Plap = f[0][2] * f[1][0] * f[2][0] + f[0][0] * f[1][2] * f[2][0] + f[0][0] * f[1][0] * f[2][2];
            v[i] = Plap * R[0] + T(2 * (qq[0] + qq[1] + qq[2]) + 3) * P * R[1] +
                   (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) * P * R[2];
        } else {
            const int a = cc == 8 ? 1 : 0, b = cc == 5 ? 1 : 2;
            const T Pa = (a == 0 ? f[0][1] : f[0][0]) * (a == 1 ? f[1][1] : f[1][0]) * f[2][0];
            const T Pb = f[0][0] * (b == 1 ? f[1][1] : f[1][0]) * (b == 2 ? f[2][1] : f[2][0]);
            const T Pab = (a == 0 ? f[0][1] : f[0][0]) * ((a == 1 || b == 1) ? f[1][1] : f[1][0]) *
                          (b == 2 ? f[2][1] : f[2][0]);
            v[i] = Pab * R[0] + (c[b] * Pa + c[a] * Pb) * R[1] + c[a] * c[b] * P * R[2];
        }
    }
}

__device__ __attribute__((always_inline)) inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __attribute__((always_inline)) inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

} // namespace kumo
