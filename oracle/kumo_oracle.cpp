// ORACLE — test infrastructure only (see kumo_oracle.hpp). CPU FP64
// restatement of the reference chem-kumo 2.1.0 evaluator; citations are to
// the reference sources. Built with g++ -std=c++20 -O3 -fopenmp
// -ffp-contract=off so that no multiply-add is fused, as in the reference's
// x86-64 builds.
#include "kumo_oracle.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <fstream>
#include <iomanip>
#include <map>
#include <memory>
#include <numbers>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Basis preparation: src/basis_math.cpp, src/basis_evaluator.cpp:15-202.

struct Power {
    int x, y, z;
};

// Gaussian FCHK order of Cartesian components (src/basis_evaluator.cpp:135-152).
std::vector<Power> cartesian_order(int l) {
    switch (l) {
    case 0: return {{0, 0, 0}};
    case 1: return {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    case 2: return {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    case 3:
        return {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0},
                {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}};
    default: break;
    }
    std::vector<Power> out;
    for (int lx = 0; lx <= l; ++lx)
        for (int ly = 0; ly <= l - lx; ++ly)
            out.push_back({lx, ly, l - lx - ly});
    return out;
}

double fact(int n) {
    double v = 1.0;
    for (long long i = 2; i <= n; ++i)
        v *= static_cast<double>(i);
    return v;
}

// Schlegel & Frisch (1995) closed form of the Cartesian->pure coefficient,
// as in src/basis_evaluator.cpp:30-96 (same published formula and factor
// grouping, so the transform rounds the same way).
double pure_coefficient(int l, int m, int lx, int ly) {
    const int lz = l - lx - ly;
    const int ma = std::abs(m);
    if (ma > l || lz < 0) return 0.0;
    const double a = std::sqrt(fact(2 * lx) * fact(2 * ly) * fact(2 * lz) * fact(l) /
                               (fact(2 * l) * fact(lx) * fact(ly) * fact(lz)));
    const double b = std::sqrt(fact(l - ma) / fact(l + ma)) / (std::pow(2.0, l) * fact(l));
    const int ilim = (l - ma) / 2;
    const int j = (l - ma - lz) / 2;
    if (2 * j != l - ma - lz) return 0.0;
    double g = 0.0;
    for (int i = 0; i <= ilim; ++i) {
        if (j < 0 || j > i) continue;
        const double c = fact(l) / (fact(i) * fact(l - i)) * fact(2 * l - 2 * i) *
                         std::pow(-1.0, i) / fact(l - ma - 2 * i) * fact(i) /
                         (fact(j) * fact(i - j));
        for (int k = 0; k <= j; ++k) {
            if (lx - 2 * k < 0 || lx - 2 * k > ma) continue;
            const double d = fact(j) / (fact(k) * fact(j - k)) * fact(ma) /
                             (fact(lx - 2 * k) * fact(ma + 2 * k - lx));
            double e = 0.0;
            if (m == 0 && lx % 2 == 0) e = std::pow(-1.0, k - lx / 2);
            const int kt = (2 * k + ma - lx) / 2;
            if (m > 0 && std::abs(ma - lx) % 2 == 0) e = std::sqrt(2.0) * std::pow(-1.0, kt);
            if (m < 0 && std::abs(ma - lx) % 2 == 1) e = std::sqrt(2.0) * std::pow(-1.0, kt);
            g += c * d * e;
        }
    }
    return a * b * g;
}

// (2l+1) x ncart, rows m = 0, +1, -1, ... (src/basis_evaluator.cpp:155-175).
std::vector<double> pure_transform(int l) {
    const auto order = cartesian_order(l);
    const int nc = static_cast<int>(order.size());
    std::vector<double> t(static_cast<std::size_t>((2 * l + 1) * nc), 0.0);
    for (int g = 0; g < nc; ++g) {
        const auto& p = order[static_cast<std::size_t>(g)];
        t[static_cast<std::size_t>(g)] = pure_coefficient(l, 0, p.x, p.y);
        for (int m = 1; m <= l; ++m) {
            t[static_cast<std::size_t>((2 * m - 1) * nc + g)] = pure_coefficient(l, m, p.x, p.y);
            t[static_cast<std::size_t>((2 * m) * nc + g)] = pure_coefficient(l, -m, p.x, p.y);
        }
    }
    return t;
}

double double_factorial_odd(int n) {
    if (n <= 0) return 1.0;
    double v = 1.0;
    for (int i = n; i >= 1; i -= 2)
        v *= static_cast<double>(i);
    return v;
}

// src/basis_math.cpp:17-25.
double primitive_normalization(double alpha, const Power& p) {
    const int l = p.x + p.y + p.z;
    const double radial = std::pow(2.0 * alpha / std::numbers::pi, 0.75);
    const double angular = std::sqrt(std::pow(4.0 * alpha, static_cast<double>(l)) /
                                     (double_factorial_odd(2 * p.x - 1) *
                                      double_factorial_odd(2 * p.y - 1) *
                                      double_factorial_odd(2 * p.z - 1)));
    return radial * angular;
}

struct Block {
    double origin[3];
    int l = 0;
    int atom = 0;
    std::vector<double> exps;
    std::vector<Power> powers;
    std::vector<double> weights; // ncart x nprim, row-major
    std::vector<double> pure;    // count x ncart, row-major; empty if Cartesian
    int offset = 0, count = 0;
    double cutoff_r2 = 0.0;
    int ncart() const { return static_cast<int>(powers.size()); }
    int nprim() const { return static_cast<int>(exps.size()); }
    double w(int k, int p) const { return weights[static_cast<std::size_t>(k * nprim() + p)]; }
};

// src/basis_evaluator.cpp:98-127: radius beyond which every AO of the block
// and its first and second derivatives are at most `cutoff`.
double shell_cutoff_radius(const Block& b, double transform_norm, double cutoff) {
    const int l = b.l;
    const int np = b.nprim();
    std::vector<double> coefficient(static_cast<std::size_t>(np));
    double start = 1.0;
    for (int p = 0; p < np; ++p) {
        const double a = b.exps[static_cast<std::size_t>(p)];
        double weight = 0.0;
        for (int k = 0; k < b.ncart(); ++k)
            weight = std::max(weight, std::abs(b.w(k, p)));
        coefficient[static_cast<std::size_t>(p)] =
            transform_norm * weight * (1.0 + l + l * l + 2.0 * a * (2.0 * l + 2.0) + 4.0 * a * a);
        start = std::max(start, std::sqrt((l + 2.0) / (2.0 * a)));
    }
    const auto bound = [&](double r) {
        double sum = 0.0;
        for (int p = 0; p < np; ++p)
            sum += coefficient[static_cast<std::size_t>(p)] * std::pow(r, l + 2) *
                   std::exp(-b.exps[static_cast<std::size_t>(p)] * r * r);
        return sum;
    };
    if (bound(start) <= cutoff) return start;
    double low = start, high = 2.0 * start;
    while (bound(high) > cutoff)
        high *= 2.0;
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (low + high);
        (bound(mid) > cutoff ? low : high) = mid;
    }
    return high;
}

constexpr double ao_cutoff = 1e-14; // include/chem/kumo/basis_evaluator.hpp:66

struct Basis {
    std::vector<Block> blocks;
    int nao = 0;
    int max_l = 0, max_cart = 0, max_prim = 0;
};

// src/basis.cpp:21-120 + src/basis_math.cpp:27-66 + src/basis_evaluator.cpp:177-202.
std::unique_ptr<Basis> make_basis(int nshell, const int* types, const int* nprim,
                                  const int* shell_atom, const double* xyz, const double* exps,
                                  const double* coef, const double* pcoef) {
    auto basis = std::make_unique<Basis>();
    std::size_t prim = 0;
    int offset = 0;
    const auto add = [&](int s, int l, const double* c, bool pure) {
        Block b;
        for (int a = 0; a < 3; ++a)
            b.origin[a] = xyz[3 * s + a];
        b.l = l;
        b.atom = shell_atom[s] - 1;
        b.exps.assign(exps + prim, exps + prim + nprim[s]);
        b.powers = cartesian_order(l);
        const int nc = b.ncart(), np = nprim[s];
        b.weights.resize(static_cast<std::size_t>(nc * np));
        for (int k = 0; k < nc; ++k)
            for (int p = 0; p < np; ++p)
                b.weights[static_cast<std::size_t>(k * np + p)] =
                    c[prim + static_cast<std::size_t>(p)] *
                    primitive_normalization(b.exps[static_cast<std::size_t>(p)],
                                            b.powers[static_cast<std::size_t>(k)]);
        if (pure && l >= 2) b.pure = pure_transform(l);
        b.offset = offset;
        b.count = b.pure.empty() ? nc : 2 * l + 1;
        offset += b.count;
        double tnorm = 1.0;
        if (!b.pure.empty()) {
            tnorm = 0.0;
            for (int r = 0; r < b.count; ++r) {
                double s2 = 0.0;
                for (int k = 0; k < nc; ++k)
                    s2 += std::abs(b.pure[static_cast<std::size_t>(r * nc + k)]);
                tnorm = std::max(tnorm, s2);
            }
        }
        const double radius = shell_cutoff_radius(b, tnorm, ao_cutoff);
        b.cutoff_r2 = radius * radius;
        basis->max_l = std::max(basis->max_l, l);
        basis->max_cart = std::max(basis->max_cart, nc);
        basis->max_prim = std::max(basis->max_prim, np);
        basis->blocks.push_back(std::move(b));
    };
    for (int s = 0; s < nshell; ++s) {
        const int t = types[s];
        if (nprim[s] <= 0) throw std::runtime_error("invalid primitive count in fchk shell");
        if (t == -1) {
            if (!pcoef) throw std::runtime_error("missing P(S=P) coefficients for sp shell");
            add(s, 0, coef, false);
            add(s, 1, pcoef, false);
        } else {
            const int l = t < 0 ? -t : t;
            add(s, l, coef, t <= -2);
        }
        prim += static_cast<std::size_t>(nprim[s]);
    }
    basis->nao = offset;
    return basis;
}

// ---------------------------------------------------------------------------
// AO evaluation: src/basis_evaluator.cpp:413-643.

inline double sqnorm(double x, double y, double z) { return (x * x + y * y) + z * z; }

struct Scratch {
    std::vector<double> radial, cart, pw, factors;
};

// AO values of one block at displacement (x, y, z), r2 = |r|^2, inside its
// cutoff (evaluate_into_impl, src/basis_evaluator.cpp:442-525): Cartesian
// component = ((x^i y^j) z^k) * sum_p w_kp exp(-a_p r^2), then the pure
// transform. Row i of the block is written to out[i * stride].
template <typename B>
void block_values(const B& b, double x, double y, double z, double r2, double* out, std::size_t stride, Scratch& s) {
    for (int p = 0; p < b.nprim(); ++p)
        s.radial[static_cast<std::size_t>(p)] = std::exp(-b.exps[static_cast<std::size_t>(p)] * r2);
    double xp[32], yp[32], zp[32];
    xp[0] = yp[0] = zp[0] = 1.0;
    for (int n = 1; n <= b.l; ++n) {
        xp[n] = xp[n - 1] * x;
        yp[n] = yp[n - 1] * y;
        zp[n] = zp[n - 1] * z;
    }
    for (int k = 0; k < b.ncart(); ++k) {
        const auto& pw = b.powers[static_cast<std::size_t>(k)];
        const double poly = (xp[pw.x] * yp[pw.y]) * zp[pw.z];
        double rs = 0.0;
        for (int p = 0; p < b.nprim(); ++p)
            rs += b.w(k, p) * s.radial[static_cast<std::size_t>(p)];
        s.cart[static_cast<std::size_t>(k)] = poly * rs;
    }
    if (b.pure.empty()) {
        for (int k = 0; k < b.ncart(); ++k)
            out[static_cast<std::size_t>(k) * stride] = s.cart[static_cast<std::size_t>(k)];
    } else {
        for (int r = 0; r < b.count; ++r) {
            double acc = 0.0;
            for (int k = 0; k < b.ncart(); ++k)
                acc += b.pure[static_cast<std::size_t>(r * b.ncart() + k)] * s.cart[static_cast<std::size_t>(k)];
            out[static_cast<std::size_t>(r) * stride] = acc;
        }
    }
}

// AO values at one point (evaluate_into_impl, src/basis_evaluator.cpp:442-525).
void ao_values_point(const Basis& B, const double* pt, double* out, Scratch& s) {
    s.radial.resize(static_cast<std::size_t>(B.max_prim));
    s.cart.resize(static_cast<std::size_t>(B.max_cart));
    for (const auto& b : B.blocks) {
        const double x = pt[0] - b.origin[0], y = pt[1] - b.origin[1], z = pt[2] - b.origin[2];
        const double r2 = sqnorm(x, y, z);
        if (r2 > b.cutoff_r2) {
            for (int i = 0; i < b.count; ++i)
                out[b.offset + i] = 0.0;
            continue;
        }
        block_values(b, x, y, z, r2, out + b.offset, 1, s);
    }
}

// Row-major panel matrix: rows = AOs, columns = panels * n (component c of
// point j at column c * n + j), the transpose of the reference's
// column-major layout with the same (row, column) meaning.
struct Panels {
    int rows = 0, cols = 0;
    std::vector<double> v;
    void resize(int r, int c) {
        rows = r;
        cols = c;
        v.assign(static_cast<std::size_t>(r) * static_cast<std::size_t>(c), 0.0);
    }
    double& at(int r, int c) { return v[static_cast<std::size_t>(r) * static_cast<std::size_t>(cols) + static_cast<std::size_t>(c)]; }
    double at(int r, int c) const { return v[static_cast<std::size_t>(r) * static_cast<std::size_t>(cols) + static_cast<std::size_t>(c)]; }
    const double* row(int r) const { return v.data() + static_cast<std::size_t>(r) * static_cast<std::size_t>(cols); }
    double* row(int r) { return v.data() + static_cast<std::size_t>(r) * static_cast<std::size_t>(cols); }
};

int panel_count(int order) { return order == 0 ? 1 : (order == 2 ? 10 : 4); }

// evaluate_batch (order 0) or evaluate_derivatives_batch (order 1, 2):
// src/basis_evaluator.cpp:512-643, 665-678.
void ao_panels(const Basis& B, int n, const double* pts, int order, Panels& out, Scratch& s) {
    const int panels = panel_count(order);
    out.resize(B.nao, panels * n);
    if (order == 0) {
        // blocks outer, points inner: the same per-(block, point) arithmetic
        // as ao_values_point, written along the rows of the panel matrix
        s.radial.resize(static_cast<std::size_t>(B.max_prim));
        s.cart.resize(static_cast<std::size_t>(B.max_cart));
        for (const auto& b : B.blocks)
            for (int j = 0; j < n; ++j) {
                const double* pt = pts + 3 * j;
                const double x = pt[0] - b.origin[0], y = pt[1] - b.origin[1], z = pt[2] - b.origin[2];
                const double r2 = sqnorm(x, y, z);
                if (r2 <= b.cutoff_r2) block_values(b, x, y, z, r2, out.row(b.offset) + j, static_cast<std::size_t>(out.cols), s);
            }
        return;
    }
    const int deg = B.max_l + 1;
    s.radial.resize(static_cast<std::size_t>(B.max_prim));
    s.pw.resize(static_cast<std::size_t>(3 * deg));
    s.factors.resize(static_cast<std::size_t>(B.max_prim * 3 * deg * 3));
    std::vector<double> cart(static_cast<std::size_t>(B.max_cart * panels));
    for (const auto& b : B.blocks) {
        const int nc = b.ncart(), np = b.nprim(), L = b.l;
        for (int j = 0; j < n; ++j) {
            const double r[3] = {pts[3 * j] - b.origin[0], pts[3 * j + 1] - b.origin[1],
                                 pts[3 * j + 2] - b.origin[2]};
            const double r2 = sqnorm(r[0], r[1], r[2]);
            if (!(r2 <= b.cutoff_r2)) continue; // zeros stay
            for (int p = 0; p < np; ++p)
                s.radial[static_cast<std::size_t>(p)] = std::exp(-b.exps[static_cast<std::size_t>(p)] * r2);
            for (int axis = 0; axis < 3; ++axis) {
                double* xp = &s.pw[static_cast<std::size_t>(axis * deg)];
                xp[0] = 1.0;
                for (int m = 1; m <= L; ++m)
                    xp[m] = xp[m - 1] * r[axis];
            }
            const auto F = [&](int p, int axis, int m) {
                return &s.factors[static_cast<std::size_t>(((p * 3 + axis) * deg + m) * 3)];
            };
            for (int p = 0; p < np; ++p) {
                const double a = b.exps[static_cast<std::size_t>(p)];
                for (int axis = 0; axis < 3; ++axis) {
                    const double* xp = &s.pw[static_cast<std::size_t>(axis * deg)];
                    const double x = r[axis];
                    for (int l = 0; l <= L; ++l) {
                        double* f = F(p, axis, l);
                        const double xl = xp[l];
                        f[0] = xl;
                        f[1] = (l ? l * xp[l - 1] : 0.0) - 2.0 * a * x * xl;
                        f[2] = order == 2 ? (l >= 2 ? l * (l - 1) * xp[l - 2] : 0.0) -
                                                2.0 * a * (2 * l + 1) * xl +
                                                4.0 * a * a * x * x * xl
                                          : 0.0;
                    }
                }
            }
            for (int k = 0; k < nc; ++k) {
                const auto& pw = b.powers[static_cast<std::size_t>(k)];
                double sums[10] = {};
                for (int p = 0; p < np; ++p) {
                    const double wgt = b.w(k, p) * s.radial[static_cast<std::size_t>(p)];
                    const double* fx = F(p, 0, pw.x);
                    const double* fy = F(p, 1, pw.y);
                    const double* fz = F(p, 2, pw.z);
                    sums[0] += wgt * fx[0] * fy[0] * fz[0];
                    sums[1] += wgt * fx[1] * fy[0] * fz[0];
                    sums[2] += wgt * fx[0] * fy[1] * fz[0];
                    sums[3] += wgt * fx[0] * fy[0] * fz[1];
                    if (order == 2) {
                        sums[4] += wgt * fx[2] * fy[0] * fz[0];
                        sums[5] += wgt * fx[1] * fy[1] * fz[0];
                        sums[6] += wgt * fx[1] * fy[0] * fz[1];
                        sums[7] += wgt * fx[0] * fy[2] * fz[0];
                        sums[8] += wgt * fx[0] * fy[1] * fz[1];
                        sums[9] += wgt * fx[0] * fy[0] * fz[2];
                    }
                }
                for (int c = 0; c < panels; ++c)
                    cart[static_cast<std::size_t>(k * panels + c)] = sums[c];
            }
            for (int c = 0; c < panels; ++c) {
                if (b.pure.empty()) {
                    for (int k = 0; k < nc; ++k)
                        out.at(b.offset + k, c * n + j) = cart[static_cast<std::size_t>(k * panels + c)];
                } else {
                    for (int row = 0; row < b.count; ++row) {
                        double acc = 0.0;
                        for (int k = 0; k < nc; ++k)
                            acc += b.pure[static_cast<std::size_t>(row * nc + k)] *
                                   cart[static_cast<std::size_t>(k * panels + c)];
                        out.at(b.offset + row, c * n + j) = acc;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Analytic overlap: src/basis_evaluator.cpp:210-278.

double ipow(double x, int n) {
    double r = 1.0;
    for (int i = 0; i < n; ++i)
        r *= x;
    return r;
}

double overlap_axis(int a, int b, double alpha, double beta, double A, double Bc) {
    const double p = alpha + beta;
    const double center = (alpha * A + beta * Bc) / p;
    const double PA = center - A, PB = center - Bc;
    const double pref = std::exp(-alpha * beta / p * (A - Bc) * (A - Bc));
    std::vector<double> mom(static_cast<std::size_t>(a + b + 1), 0.0);
    mom[0] = std::sqrt(std::numbers::pi / p);
    for (int k = 2; k <= a + b; k += 2)
        mom[static_cast<std::size_t>(k)] = (k - 1) * mom[static_cast<std::size_t>(k - 2)] / (2.0 * p);
    double res = 0.0, ba = 1.0;
    for (int i = 0; i <= a; ++i) {
        double bb = 1.0;
        for (int j = 0; j <= b; ++j) {
            res += ba * bb * ipow(PA, a - i) * ipow(PB, b - j) * mom[static_cast<std::size_t>(i + j)];
            if (j < b) bb *= static_cast<double>(b - j) / (j + 1);
        }
        if (i < a) ba *= static_cast<double>(a - i) / (i + 1);
    }
    return pref * res;
}

std::vector<double> overlap(const Basis& B) {
    const int n = B.nao;
    std::vector<double> S(static_cast<std::size_t>(n) * static_cast<std::size_t>(n), 0.0);
    for (std::size_t ia = 0; ia < B.blocks.size(); ++ia) {
        const auto& A = B.blocks[ia];
        for (std::size_t ib = ia; ib < B.blocks.size(); ++ib) {
            const auto& Bb = B.blocks[ib];
            const int na = A.ncart(), nb = Bb.ncart();
            std::vector<double> cart(static_cast<std::size_t>(na * nb), 0.0);
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    const auto& x = A.powers[static_cast<std::size_t>(i)];
                    const auto& y = Bb.powers[static_cast<std::size_t>(j)];
                    double v = 0.0;
                    for (int p = 0; p < A.nprim(); ++p)
                        for (int q = 0; q < Bb.nprim(); ++q) {
                            const double al = A.exps[static_cast<std::size_t>(p)], be = Bb.exps[static_cast<std::size_t>(q)];
                            const double prim = overlap_axis(x.x, y.x, al, be, A.origin[0], Bb.origin[0]) *
                                                overlap_axis(x.y, y.y, al, be, A.origin[1], Bb.origin[1]) *
                                                overlap_axis(x.z, y.z, al, be, A.origin[2], Bb.origin[2]);
                            v += A.w(i, p) * Bb.w(j, q) * prim;
                        }
                    cart[static_cast<std::size_t>(i * nb + j)] = v;
                }
            // block = TA * cart * TB^T
            std::vector<double> left(static_cast<std::size_t>(A.count * nb));
            for (int r = 0; r < A.count; ++r)
                for (int j = 0; j < nb; ++j) {
                    if (A.pure.empty()) {
                        left[static_cast<std::size_t>(r * nb + j)] = cart[static_cast<std::size_t>(r * nb + j)];
                    } else {
                        double acc = 0.0;
                        for (int k = 0; k < na; ++k)
                            acc += A.pure[static_cast<std::size_t>(r * na + k)] * cart[static_cast<std::size_t>(k * nb + j)];
                        left[static_cast<std::size_t>(r * nb + j)] = acc;
                    }
                }
            for (int r = 0; r < A.count; ++r)
                for (int c = 0; c < Bb.count; ++c) {
                    double v;
                    if (Bb.pure.empty()) {
                        v = left[static_cast<std::size_t>(r * nb + c)];
                    } else {
                        v = 0.0;
                        for (int k = 0; k < nb; ++k)
                            v += left[static_cast<std::size_t>(r * nb + k)] * Bb.pure[static_cast<std::size_t>(c * nb + k)];
                    }
                    const std::size_t i = static_cast<std::size_t>(A.offset + r), j = static_cast<std::size_t>(Bb.offset + c);
                    S[i * static_cast<std::size_t>(n) + j] = v;
                    S[j * static_cast<std::size_t>(n) + i] = v;
                }
        }
    }
    return S;
}

// ---------------------------------------------------------------------------
// Molecule: src/molecule.cpp, src/backend_selection.hpp:43-78.

struct Atom {
    int z;
    double charge;
    double pos[3];
};

struct Molecule {
    const Basis* basis = nullptr;
    std::vector<Atom> atoms;
    int n = 0;
    std::vector<double> P, S; // total, spin (row-major n x n)
    std::vector<double> Ca, Cb; // n x nmo, row-major (AO rows, MO columns)
    std::vector<double> Ea, Eb; // orbital energies
    int nmo_a = 0, nmo_b = 0, nalpha = 0, nbeta = 0;
    bool unrestricted = false;
    bool eligible = false;
    const std::vector<double>& C(int channel) const { return (channel == 1 && unrestricted) ? Cb : Ca; }
    int nmo(int channel) const { return (channel == 1 && unrestricted) ? nmo_b : nmo_a; }
};

double frobenius(const std::vector<double>& m) {
    double s = 0.0;
    for (double v : m)
        s += v * v;
    return std::sqrt(s);
}

// validate_orbital_backend (src/backend_selection.hpp:43-78).
bool check_eligible(const Molecule& m) {
    const int n = m.n;
    const auto& cb = m.unrestricted ? m.Cb : m.Ca;
    const int nmob = m.unrestricted ? m.nmo_b : m.nmo_a;
    if (m.nalpha > m.nmo_a || m.nbeta > nmob || (!m.unrestricted && m.nalpha != m.nbeta)) return false;
    if (m.Ca.empty()) return false;
    std::vector<double> pa(static_cast<std::size_t>(n * n), 0.0), pb(static_cast<std::size_t>(n * n), 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double sa = 0.0, sb = 0.0;
            for (int k = 0; k < m.nalpha; ++k)
                sa += m.Ca[static_cast<std::size_t>(i * m.nmo_a + k)] * m.Ca[static_cast<std::size_t>(j * m.nmo_a + k)];
            for (int k = 0; k < m.nbeta; ++k)
                sb += cb[static_cast<std::size_t>(i * nmob + k)] * cb[static_cast<std::size_t>(j * nmob + k)];
            pa[static_cast<std::size_t>(i * n + j)] = sa;
            pb[static_cast<std::size_t>(i * n + j)] = sb;
        }
    std::vector<double> rt(pa.size()), rs(pa.size()), et(pa.size()), es(pa.size());
    for (std::size_t q = 0; q < pa.size(); ++q) {
        rt[q] = pa[q] + pb[q];
        rs[q] = pa[q] - pb[q];
        et[q] = m.P[q] - rt[q];
        es[q] = m.S[q] - rs[q];
    }
    const double te = frobenius(et), se = frobenius(es);
    const double ts = std::max(frobenius(m.P), frobenius(rt));
    const double ss = std::max({frobenius(m.S), frobenius(rs), ts});
    return te <= 1e-10 + 1e-8 * ts && se <= 1e-10 + 1e-8 * ss;
}

// ---------------------------------------------------------------------------
// Request plan: src/evaluation_plan.hpp:37-273, src/automatic_backend.hpp:38-50.

enum Kind {
    Density = 0, Gradient, Hessian, Laplacian, Tau, Orbital, Elf, ReducedGradient,
    SignLambda2Density, ElectrostaticPotential, Lol, InteractionRegionIndicator,
    DensityOverlapRegionsIndicator, HamiltonianKineticEnergy, PotentialEnergyDensity,
    EnergyDensity, SpinPolarization, ShannonEntropyDensity, FisherInformationDensity,
    AverageLocalIonizationEnergy, LocalElectronAffinity, LocalElectronAttachmentEnergy,
    LocalElectronegativity, LocalHardness, OrbitalGradient, OrbitalHessian,
    OrbitalWeightedFukuiPlus, OrbitalWeightedFukuiMinus, OrbitalWeightedFukuiZero,
    OrbitalWeightedDualDescriptor, TauComponents, HamiltonianKineticEnergyComponents,
    PotentialKineticRatio, EnergyDensityPerElectron, KineticEnergyPerElectron,
    ElfTsirelsonStash, LolTsirelsonStash, KindCount
};
enum Comp { Total = 0, Alpha = 1, Beta = 2, Difference = 3 };
enum Backend { AUTO = 0, AO_P = 1, AO_MO = 2 };

int channels(int k) {
    switch (k) {
    case Gradient: case OrbitalGradient: case TauComponents: case HamiltonianKineticEnergyComponents: return 3;
    case Hessian: case OrbitalHessian: return 6;
    default: return 1;
    }
}

bool is_orbital_property(int k) { return k == Orbital || k == OrbitalGradient || k == OrbitalHessian; }
bool is_orbital_energy(int k) {
    return (k >= AverageLocalIonizationEnergy && k <= LocalHardness) ||
           (k >= OrbitalWeightedFukuiPlus && k <= OrbitalWeightedDualDescriptor);
}

struct ComponentPlan {
    bool density = false, gradient = false, hessian = false, laplacian = false, tau = false;
    bool elf = false, elf_input = false, reduced_gradient = false, sign_lambda2 = false;
    bool lol = false, lol_input = false, ge_input = false, gradient_input = false;
    bool hessian_input = false, energy_input = false, tau_input = false, value_input = false;
    bool esp = false;
    bool derivatives() const {
        return gradient || hessian || laplacian || tau || elf_input || reduced_gradient || sign_lambda2 ||
               lol_input || gradient_input || hessian_input || energy_input || tau_input || ge_input ||
               value_input;
    }
    int ao_order() const {
        if (hessian || laplacian || sign_lambda2 || hessian_input || energy_input || ge_input) return 2;
        if (gradient || tau || elf_input || reduced_gradient || lol_input || gradient_input || tau_input || value_input)
            return 1;
        return density ? 0 : -1;
    }
};

struct Plan {
    std::array<ComponentPlan, 4> c{};
    bool orbitals = false, orbital_energies = false;
    bool virtual_orbitals = false, orbital_weighted_fukui = false;
    int max_order = -1;
};

// orbital_energy_name (src/evaluation_plan.hpp:112-128).
const char* orbital_energy_name(int kind) {
    switch (kind) {
    case AverageLocalIonizationEnergy: return "the average local ionization energy";
    case LocalElectronAffinity: return "the local electron affinity";
    case LocalElectronAttachmentEnergy: return "the local electron attachment energy";
    case LocalElectronegativity: return "the local electronegativity";
    case LocalHardness: return "the local hardness";
    case OrbitalWeightedFukuiPlus: return "the orbital-weighted Fukui function f+";
    case OrbitalWeightedFukuiMinus: return "the orbital-weighted Fukui function f-";
    case OrbitalWeightedFukuiZero: return "the orbital-weighted Fukui function f0";
    default: return "the orbital-weighted dual descriptor";
    }
}

void require_total(const orc_property& p, const char* what) {
    if (p.component != Total) throw std::invalid_argument(std::string(what) + " is defined only for the total density");
}

void spin_inputs(Plan& plan, const orc_property& p, bool ComponentPlan::*input, const char* what) {
    if (p.component == Difference) throw std::invalid_argument(std::string(what) + " is not defined for the spin-difference density");
    if (p.component == Total) {
        plan.c[Alpha].*input = true;
        plan.c[Beta].*input = true;
    } else {
        plan.c[static_cast<std::size_t>(p.component)].*input = true;
    }
}

Plan make_plan(const std::vector<orc_property>& props) {
    Plan plan;
    for (const auto& p : props) {
        if (p.component < 0 || p.component > 3) throw std::invalid_argument("invalid density spin component");
        if (is_orbital_property(p.kind)) {
            if (p.orbital_index == 0) throw std::invalid_argument("orbital indices are one-based");
            plan.orbitals = true;
            plan.max_order = std::max(plan.max_order, 0);
            continue;
        }
        auto& c = plan.c[static_cast<std::size_t>(p.component)];
        switch (p.kind) {
        case Density: c.density = true; break;
        case Gradient: c.gradient = true; break;
        case Hessian: c.hessian = true; break;
        case Laplacian: c.laplacian = true; break;
        case Tau: c.tau = true; break;
        case Elf: c.elf = true; spin_inputs(plan, p, &ComponentPlan::elf_input, "ELF"); break;
        case ReducedGradient: require_total(p, "the reduced density gradient"); c.reduced_gradient = true; break;
        case SignLambda2Density: require_total(p, "sign(lambda2) rho"); c.sign_lambda2 = true; break;
        case ElectrostaticPotential: require_total(p, "the electrostatic potential"); c.esp = true; break;
        case Lol: c.lol = true; spin_inputs(plan, p, &ComponentPlan::lol_input, "LOL"); break;
        case ElfTsirelsonStash: spin_inputs(plan, p, &ComponentPlan::ge_input, "the Tsirelson-Stash ELF"); break;
        case LolTsirelsonStash: spin_inputs(plan, p, &ComponentPlan::ge_input, "the Tsirelson-Stash LOL"); break;
        case InteractionRegionIndicator: require_total(p, "the interaction region indicator"); c.gradient_input = true; break;
        case ShannonEntropyDensity: require_total(p, "the Shannon entropy density"); c.gradient_input = true; break;
        case FisherInformationDensity: require_total(p, "the Fisher information density"); c.gradient_input = true; break;
        case DensityOverlapRegionsIndicator: require_total(p, "the density overlap regions indicator"); c.hessian_input = true; break;
        case HamiltonianKineticEnergy: c.energy_input = true; break;
        case PotentialEnergyDensity: require_total(p, "the potential energy density"); c.energy_input = true; break;
        case EnergyDensity: require_total(p, "the energy density"); c.energy_input = true; break;
        case HamiltonianKineticEnergyComponents: c.energy_input = true; break;
        case TauComponents: c.tau_input = true; break;
        case PotentialKineticRatio: require_total(p, "|V| / G"); c.energy_input = true; break;
        case EnergyDensityPerElectron: require_total(p, "the energy density per electron"); c.energy_input = true; break;
        case KineticEnergyPerElectron: require_total(p, "the kinetic energy per electron"); c.tau_input = true; break;
        case SpinPolarization:
            if (p.component != Difference) throw std::invalid_argument("the spin polarization is defined only for the spin component");
            plan.c[Total].value_input = true;
            c.value_input = true;
            break;
        case AverageLocalIonizationEnergy:
            require_total(p, orbital_energy_name(p.kind));
            plan.orbital_energies = true;
            break;
        case LocalElectronAffinity: case LocalElectronAttachmentEnergy: case LocalElectronegativity:
        case LocalHardness:
            require_total(p, orbital_energy_name(p.kind));
            plan.orbital_energies = plan.virtual_orbitals = true;
            break;
        case OrbitalWeightedFukuiPlus: case OrbitalWeightedFukuiMinus: case OrbitalWeightedFukuiZero:
        case OrbitalWeightedDualDescriptor:
            require_total(p, orbital_energy_name(p.kind));
            plan.orbital_energies = plan.orbital_weighted_fukui = true;
            break;
        default:
            throw std::invalid_argument("unknown property kind");
        }
    }
    for (const auto& c : plan.c)
        plan.max_order = std::max(plan.max_order, c.ao_order());
    return plan;
}

int choose_backend(int requested, bool eligible, const Plan& plan) {
    if (requested == AO_P) return AO_P;
    if (requested == AO_MO) {
        if (!eligible) throw std::invalid_argument("AO_MO requested for ineligible orbital representation");
        return AO_MO;
    }
    if (requested != AUTO) throw std::invalid_argument("unknown evaluation backend");
    if (!eligible) return AO_P;
    for (const auto& c : plan.c)
        if (c.ao_order() >= 0) return AO_MO;
    return AO_P;
}

// ---------------------------------------------------------------------------
// Density kernels: src/density.cpp, src/ao_screening.hpp.

struct DD {
    double value = 0.0;
    double gradient[3] = {0, 0, 0};
    double hessian[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double laplacian = 0.0;
    double tau = 0.0;
    double tau_axes[3] = {0, 0, 0};
};

constexpr int hcol[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}}; // src/derivative_layout.hpp:13

double trace(const DD& d) { return d.hessian[0][0] + d.hessian[1][1] + d.hessian[2][2]; }

// src/density.cpp:50-60.
DD combine(const DD& a, const DD& b, double sign, double scale) {
    DD d;
    d.value = scale * (a.value + sign * b.value);
    for (int i = 0; i < 3; ++i) {
        d.gradient[i] = scale * (a.gradient[i] + sign * b.gradient[i]);
        d.tau_axes[i] = scale * (a.tau_axes[i] + sign * b.tau_axes[i]);
        for (int j = 0; j < 3; ++j)
            d.hessian[i][j] = scale * (a.hessian[i][j] + sign * b.hessian[i][j]);
    }
    d.laplacian = trace(d);
    d.tau = scale * (a.tau + sign * b.tau);
    return d;
}

// significant_aos (src/ao_screening.hpp:35-50).
std::vector<int> significant(const Panels& ao, double threshold) {
    std::vector<int> rows;
    if (threshold == 0.0 || ao.cols == 0) {
        for (int r = 0; r < ao.rows; ++r)
            rows.push_back(r);
        return rows;
    }
    for (int r = 0; r < ao.rows; ++r) {
        double mx = 0.0;
        const double* row = ao.row(r);
        for (int c = 0; c < ao.cols; ++c)
            mx = std::max(mx, std::abs(row[c]));
        if (mx > threshold) rows.push_back(r);
    }
    return rows;
}

// out(i, col) = sum_k A(kept[k], i) * X(kept[k] row, col), A row-major n x lda.
// Blocked over 4 output rows; the inner loop runs along contiguous columns.
// out[i][c] = sum over k in kept (in list order) of A[k][i] * X[k][col0 + c]:
// register tiles of 4 rows x 8 columns (two 4-wide vectors per row) whose
// accumulators stay in registers over the kept list; each output element is
// summed in kept order from zero, with separate multiply and add (no fused
// operations), so the result equals the plain triple loop bit for bit.
typedef double v4d __attribute__((vector_size(32)));

void project(const double* A, int lda, int m_out, const std::vector<int>& kept, const Panels& X,
             int col0, int ncols, std::vector<double>& out) {
    out.assign(static_cast<std::size_t>(m_out) * static_cast<std::size_t>(ncols), 0.0);
    const std::size_t nk = kept.size();
    const int c8 = ncols / 8 * 8;
    // the four coefficients of each kept row, packed once per row group
    std::vector<double> pack(4 * nk);
    int i = 0;
    for (; i + 4 <= m_out; i += 4) {
        double* o = out.data() + static_cast<std::size_t>(i) * ncols;
        for (std::size_t q = 0; q < nk; ++q)
            for (int r = 0; r < 4; ++r)
                pack[4 * q + static_cast<std::size_t>(r)] = A[static_cast<std::size_t>(kept[q]) * lda + i + r];
        for (int c = 0; c < c8; c += 8) {
            v4d acc[4][2] = {};
            for (std::size_t q = 0; q < nk; ++q) {
                const int k = kept[q];
                const double* a = pack.data() + 4 * q;
                const double* x = X.row(k) + col0 + c;
                v4d x0, x1;
                std::memcpy(&x0, x, sizeof(x0));
                std::memcpy(&x1, x + 4, sizeof(x1));
                for (int r = 0; r < 4; ++r) {
                    const v4d ar = {a[r], a[r], a[r], a[r]};
                    acc[r][0] += ar * x0;
                    acc[r][1] += ar * x1;
                }
            }
            for (int r = 0; r < 4; ++r) {
                std::memcpy(o + static_cast<std::size_t>(r) * ncols + c, &acc[r][0], sizeof(v4d));
                std::memcpy(o + static_cast<std::size_t>(r) * ncols + c + 4, &acc[r][1], sizeof(v4d));
            }
        }
        for (int c = c8; c < ncols; ++c)
            for (int r = 0; r < 4; ++r) {
                double sum = 0.0;
                for (std::size_t q = 0; q < nk; ++q) {
                    const int k = kept[q];
                    sum += A[static_cast<std::size_t>(k) * lda + i + r] * X.row(k)[col0 + c];
                }
                o[static_cast<std::size_t>(r) * ncols + c] = sum;
            }
    }
    for (; i < m_out; ++i) {
        double* o = out.data() + static_cast<std::size_t>(i) * ncols;
        for (int k : kept) {
            const double c0 = A[static_cast<std::size_t>(k) * lda + i];
            const double* x = X.row(k) + col0;
            for (int c = 0; c < ncols; ++c)
                o[c] += c0 * x[c];
        }
    }
}

using BlockDD = std::map<int, std::vector<DD>>;

// density_derivatives_mo (src/density.cpp:234-351), one block of b points.
BlockDD density_mo(const Molecule& M, const double* pts, int b, const std::vector<int>& comps, int ao_order,
                   double threshold, Scratch& s) {
    BlockDD result;
    for (int c : comps)
        result[c].assign(static_cast<std::size_t>(b), DD{});
    if (result.empty() || b == 0) return result;
    const bool need_a = result.count(Alpha) || result.count(Total) || result.count(Difference);
    const bool need_b = result.count(Beta) || result.count(Total) || result.count(Difference);
    const bool beta_is_alpha = need_a && need_b && !M.unrestricted && M.nalpha == M.nbeta;
    const int panels = panel_count(ao_order);
    Panels ao;
    ao_panels(*M.basis, b, pts, ao_order, ao, s);
    const auto kept = significant(ao, threshold);
    // Psi[panel] = C_occ,kept^T * Phi_panel,kept, rows = occupied MOs.
    const auto projectall = [&](int channel, int nocc, std::vector<std::vector<double>>& psi) {
        psi.assign(static_cast<std::size_t>(panels), {});
        if (nocc == 0) return;
        const auto& C = M.C(channel);
        const int nmo = M.nmo(channel);
        for (int a = 0; a < panels; ++a)
            project(C.data(), nmo, nocc, kept, ao, a * b, b, psi[static_cast<std::size_t>(a)]);
    };
    std::vector<std::vector<double>> va, vb;
    if (need_a) projectall(0, M.nalpha, va);
    if (need_b && !beta_is_alpha) projectall(1, M.nbeta, vb);
    const auto contract = [&](const std::vector<std::vector<double>>& v, int nocc, int j) {
        DD d;
        if (nocc == 0) return d;
        const auto dot = [&](int p, int q) {
            double sum = 0.0;
            for (int i = 0; i < nocc; ++i)
                sum += v[static_cast<std::size_t>(p)][static_cast<std::size_t>(i * b + j)] *
                       v[static_cast<std::size_t>(q)][static_cast<std::size_t>(i * b + j)];
            return sum;
        };
        d.value = dot(0, 0);
        if (ao_order == 0) return d;
        for (int a = 0; a < 3; ++a) {
            d.gradient[a] = 2.0 * dot(0, a + 1);
            d.tau_axes[a] = 0.5 * dot(a + 1, a + 1);
            d.tau += d.tau_axes[a];
            if (ao_order == 2)
                for (int c = a; c < 3; ++c) {
                    const double v2 = 2.0 * (dot(0, 4 + hcol[a][c]) + dot(a + 1, c + 1));
                    d.hessian[a][c] = d.hessian[c][a] = v2;
                }
        }
        d.laplacian = trace(d);
        return d;
    };
    for (int j = 0; j < b; ++j) {
        const DD alpha = need_a ? contract(va, M.nalpha, j) : DD{};
        DD beta;
        if (beta_is_alpha) beta = alpha;
        else if (need_b) beta = contract(vb, M.nbeta, j);
        const auto put = [&](int c, const DD& d) {
            auto it = result.find(c);
            if (it != result.end()) it->second[static_cast<std::size_t>(j)] = d;
        };
        put(Alpha, alpha);
        put(Beta, beta);
        if (result.count(Total)) put(Total, combine(alpha, beta, 1.0, 1.0));
        if (result.count(Difference)) put(Difference, combine(alpha, beta, -1.0, 1.0));
    }
    return result;
}

// contract_derivatives (src/density.cpp:26-44).
void contract_p(const Panels& ao, const std::vector<int>& kept, const std::vector<double>& prod, int b, int j,
                int ao_order, DD& out) {
    // prod: m x 4b (row i = kept row i); ao rows indexed through kept.
    const int m = static_cast<int>(kept.size());
    const auto dot = [&](int panel, int ppanel) {
        double s = 0.0;
        for (int i = 0; i < m; ++i)
            s += ao.at(kept[static_cast<std::size_t>(i)], panel * b + j) * prod[static_cast<std::size_t>(i) * static_cast<std::size_t>(4 * b) + static_cast<std::size_t>(ppanel * b + j)];
        return s;
    };
    out.value = dot(0, 0);
    for (int a = 0; a < 3; ++a) {
        out.gradient[a] = 2.0 * dot(a + 1, 0);
        out.tau_axes[a] = 0.5 * dot(a + 1, a + 1);
        out.tau += out.tau_axes[a];
        if (ao_order == 2)
            for (int c = a; c < 3; ++c) {
                const double v = 2.0 * (dot(4 + hcol[a][c], 0) + dot(a + 1, c + 1));
                out.hessian[a][c] = out.hessian[c][a] = v;
            }
    }
    out.laplacian = trace(out);
}

// W = P_kept * ao_kept[:, 0:4b] (row i of W = kept AO i).
void p_product(const std::vector<double>& P, int n, const std::vector<int>& kept, const Panels& ao, int ncols,
               std::vector<double>& out) {
    const int m = static_cast<int>(kept.size());
    out.assign(static_cast<std::size_t>(m) * static_cast<std::size_t>(ncols), 0.0);
    for (int i = 0; i < m; ++i) {
        double* o = out.data() + static_cast<std::size_t>(i) * ncols;
        const double* prow = P.data() + static_cast<std::size_t>(kept[static_cast<std::size_t>(i)]) * n;
        for (int k : kept) {
            const double c = prow[k];
            const double* x = ao.row(k);
            for (int col = 0; col < ncols; ++col)
                o[col] += c * x[col];
        }
    }
}

// density_derivatives_shared (src/density.cpp:149-229), one block.
BlockDD density_p(const Molecule& M, const double* pts, int b, const std::vector<int>& comps, int ao_order,
                  double threshold, Scratch& s) {
    BlockDD result;
    bool tn = false, sn = false;
    for (int c : comps) {
        if (c == Total) tn = true;
        else if (c == Difference) sn = true;
        else tn = sn = true;
        result[c].assign(static_cast<std::size_t>(b), DD{});
    }
    if (result.empty() || b == 0) return result;
    Panels ao;
    ao_panels(*M.basis, b, pts, ao_order, ao, s);
    const auto kept = significant(ao, threshold);
    std::vector<double> pt, ps;
    if (tn) p_product(M.P, M.n, kept, ao, 4 * b, pt);
    if (sn) p_product(M.S, M.n, kept, ao, 4 * b, ps);
    for (int j = 0; j < b; ++j) {
        DD t, sp;
        if (tn) contract_p(ao, kept, pt, b, j, ao_order, t);
        if (sn) contract_p(ao, kept, ps, b, j, ao_order, sp);
        const auto put = [&](int c, const DD& d) {
            auto it = result.find(c);
            if (it != result.end()) it->second[static_cast<std::size_t>(j)] = d;
        };
        put(Total, t);
        put(Difference, sp);
        if (result.count(Alpha)) put(Alpha, combine(t, sp, 1.0, 0.5));
        if (result.count(Beta)) put(Beta, combine(t, sp, -1.0, 0.5));
    }
    return result;
}

// ---------------------------------------------------------------------------
// Derived fields: src/evaluator.cpp:127-353.

const double kFermi = 0.3 * std::pow(6.0 * std::numbers::pi * std::numbers::pi, 2.0 / 3.0);

double gnorm2(const DD& d) { return (d.gradient[0] * d.gradient[0] + d.gradient[1] * d.gradient[1]) + d.gradient[2] * d.gradient[2]; }

double elf(std::initializer_list<const DD*> ch) {
    double pauli = 0.0, uniform = 0.0;
    for (const DD* d : ch) {
        if (!(d->value > 0.0)) continue;
        pauli += d->tau - gnorm2(*d) / (8.0 * d->value);
        uniform += kFermi * std::pow(d->value, 5.0 / 3.0);
    }
    if (!(uniform > 0.0)) return 0.0;
    const double chi = pauli / uniform;
    return 1.0 / (1.0 + chi * chi);
}

double reduced_gradient(const DD& d) {
    const double scale = 2.0 * std::cbrt(3.0 * std::numbers::pi * std::numbers::pi);
    if (!(d.value > 0.0)) return 100.0;
    const double den = scale * std::pow(d.value, 4.0 / 3.0);
    const double norm = std::sqrt(gnorm2(d));
    if (!(den > 0.0) || !(norm < 100.0 * den)) return 100.0;
    return norm / den;
}

// Middle eigenvalue of the symmetric 3x3 Hessian by cyclic Jacobi rotations.
double middle_eigenvalue(const double h[3][3]) {
    double a[3][3];
    std::memcpy(a, h, sizeof(a));
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
    }
    double e[3] = {a[0][0], a[1][1], a[2][2]};
    std::sort(e, e + 3);
    return e[1];
}

double sign_lambda2(const DD& d) { return middle_eigenvalue(d.hessian) < 0.0 ? -d.value : d.value; }

double lol(std::initializer_list<const DD*> ch) {
    double tau = 0.0, uniform = 0.0;
    for (const DD* d : ch) {
        tau += d->tau;
        if (d->value > 0.0) uniform += kFermi * std::pow(d->value, 5.0 / 3.0);
    }
    if (!(uniform > 0.0)) return 0.0;
    if (!(tau > 0.0)) return 1.0;
    return uniform / (tau + uniform);
}

double ge_locator(std::initializer_list<const DD*> ch, bool is_lol) {
    double uniform = 0.0, pauli = 0.0, kinetic = 0.0;
    for (const DD* d : ch) {
        if (!(d->value > 0.0)) continue;
        const double ref = kFermi * std::pow(d->value, 5.0 / 3.0);
        const double wz = gnorm2(*d) / d->value;
        uniform += ref;
        pauli += ref - wz / 9.0 + d->laplacian / 6.0;
        kinetic += ref + wz / 72.0 + d->laplacian / 6.0;
    }
    if (!(uniform > 0.0)) return 0.0;
    if (is_lol) {
        const double den = kinetic + uniform;
        return den != 0.0 ? uniform / den : 0.0;
    }
    const double chi = pauli / uniform;
    return 1.0 / (1.0 + chi * chi);
}

double iri(const DD& d) {
    if (!(d.value > 0.0)) return 100.0;
    const double v = std::sqrt(gnorm2(d)) / std::pow(d.value, 1.1);
    return v < 100.0 ? v : 100.0;
}

double dori(const DD& d) {
    if (!(d.value > 0.0)) return 0.0;
    const double q[3] = {d.gradient[0] / d.value, d.gradient[1] / d.value, d.gradient[2] / d.value};
    const double q2 = (q[0] * q[0] + q[1] * q[1]) + q[2] * q[2];
    if (!(q2 > 0.0)) return 1.0;
    double w[3];
    for (int a = 0; a < 3; ++a)
        w[a] = (d.hessian[a][0] * q[0] + d.hessian[a][1] * q[1] + d.hessian[a][2] * q[2]) / d.value - q2 * q[a];
    const double theta = 4.0 * ((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]) / (q2 * q2 * q2);
    if (!std::isfinite(theta)) return 1.0;
    return theta / (1.0 + theta);
}

// store_block (src/evaluator.cpp:248-353).
void store_block(int kind, int component, const BlockDD& block, int count, std::int64_t start, double** out) {
    if ((kind == Elf || kind == Lol || kind == ElfTsirelsonStash || kind == LolTsirelsonStash) && component == Total) {
        const auto& a = block.at(Alpha);
        const auto& b = block.at(Beta);
        for (int l = 0; l < count; ++l) {
            const DD* pa = &a[static_cast<std::size_t>(l)];
            const DD* pb = &b[static_cast<std::size_t>(l)];
            double v;
            switch (kind) {
            case Elf: v = elf({pa, pb}); break;
            case Lol: v = lol({pa, pb}); break;
            case ElfTsirelsonStash: v = ge_locator({pa, pb}, false); break;
            default: v = ge_locator({pa, pb}, true); break;
            }
            out[0][start + l] = v;
        }
        return;
    }
    if (kind == SpinPolarization) {
        const auto& t = block.at(Total);
        const auto& s = block.at(Difference);
        for (int l = 0; l < count; ++l)
            out[0][start + l] = t[static_cast<std::size_t>(l)].value > 0.0 ? s[static_cast<std::size_t>(l)].value / t[static_cast<std::size_t>(l)].value : 0.0;
        return;
    }
    const auto& dv = block.at(component);
    for (int l = 0; l < count; ++l) {
        const DD& d = dv[static_cast<std::size_t>(l)];
        const std::int64_t p = start + l;
        switch (kind) {
        case Density: out[0][p] = d.value; break;
        case Gradient: for (int a = 0; a < 3; ++a) out[a][p] = d.gradient[a]; break;
        case Hessian:
            out[0][p] = d.hessian[0][0]; out[1][p] = d.hessian[0][1]; out[2][p] = d.hessian[0][2];
            out[3][p] = d.hessian[1][1]; out[4][p] = d.hessian[1][2]; out[5][p] = d.hessian[2][2];
            break;
        case Laplacian: out[0][p] = d.laplacian; break;
        case Tau: out[0][p] = d.tau; break;
        case Elf: out[0][p] = elf({&d}); break;
        case ReducedGradient: out[0][p] = reduced_gradient(d); break;
        case SignLambda2Density: out[0][p] = sign_lambda2(d); break;
        case Lol: out[0][p] = lol({&d}); break;
        case ElfTsirelsonStash: out[0][p] = ge_locator({&d}, false); break;
        case LolTsirelsonStash: out[0][p] = ge_locator({&d}, true); break;
        case InteractionRegionIndicator: out[0][p] = iri(d); break;
        case DensityOverlapRegionsIndicator: out[0][p] = dori(d); break;
        case HamiltonianKineticEnergy: out[0][p] = d.tau - 0.25 * d.laplacian; break;
        case PotentialEnergyDensity: out[0][p] = 0.25 * d.laplacian - 2.0 * d.tau; break;
        case EnergyDensity: out[0][p] = 0.25 * d.laplacian - d.tau; break;
        case TauComponents: for (int a = 0; a < 3; ++a) out[a][p] = d.tau_axes[a]; break;
        case HamiltonianKineticEnergyComponents:
            for (int a = 0; a < 3; ++a) out[a][p] = d.tau_axes[a] - 0.25 * d.hessian[a][a];
            break;
        case PotentialKineticRatio: out[0][p] = d.tau > 0.0 ? std::abs(0.25 * d.laplacian - 2.0 * d.tau) / d.tau : 0.0; break;
        case EnergyDensityPerElectron: out[0][p] = d.value > 0.0 ? (0.25 * d.laplacian - d.tau) / d.value : 0.0; break;
        case KineticEnergyPerElectron: out[0][p] = d.value > 0.0 ? d.tau / d.value : 0.0; break;
        case ShannonEntropyDensity: out[0][p] = d.value > 0.0 ? -d.value * std::log(d.value) : 0.0; break;
        case FisherInformationDensity: out[0][p] = d.value > 0.0 ? gnorm2(d) / d.value : 0.0; break;
        default: throw std::logic_error("invalid derivative property");
        }
    }
}

// ---------------------------------------------------------------------------
// Boys function: src/boys_function.cpp (table + 7-term Taylor below T = 40,
// closed form and upward recursion above; the same published scheme and
// arithmetic order, so that it matches the reference bit for bit).

constexpr int kBoysMax = 32, kTaylor = 7, kOrders = kBoysMax + kTaylor, kPoints = 801;
constexpr double kSpacing = 0.05, kEnd = 40.0;

const std::vector<double>& boys_table() {
    static const std::vector<double> table = [] {
        std::vector<double> t(static_cast<std::size_t>(kPoints * kOrders));
        constexpr int top = kOrders - 1;
        for (int i = 0; i < kPoints; ++i) {
            const double T = i * kSpacing;
            const double decay = std::exp(-T);
            double term = 1.0 / (2.0 * top + 1.0), sum = term;
            for (int k = 1; k < 10000 && term > 1e-18 * sum; ++k) {
                term *= 2.0 * T / (2.0 * top + 2.0 * k + 1.0);
                sum += term;
            }
            double* row = t.data() + static_cast<std::size_t>(i * kOrders);
            row[top] = decay * sum;
            for (int m = top; m > 0; --m)
                row[m - 1] = (2.0 * T * row[m] + decay) / (2.0 * m - 1.0);
        }
        return t;
    }();
    return table;
}

void boys(double T, int top, double* values) {
    const double decay = std::exp(-T);
    if (T >= kEnd) {
        values[0] = 0.5 * std::sqrt(std::numbers::pi / T);
        for (int m = 0; m < top; ++m)
            values[m + 1] = ((2.0 * m + 1.0) * values[m] - decay) / (2.0 * T);
        return;
    }
    const int point = static_cast<int>(std::lround(T / kSpacing));
    const double delta = T - point * kSpacing;
    const double* row = boys_table().data() + static_cast<std::size_t>(point * kOrders);
    double sum = 0.0, factor = 1.0;
    for (int k = 0; k < kTaylor; ++k) {
        sum += row[top + k] * factor;
        factor *= -delta / (k + 1.0);
    }
    values[top] = sum;
    for (int m = top; m > 0; --m)
        values[m - 1] = (2.0 * T * values[m] + decay) / (2.0 * m - 1.0);
}

// ---------------------------------------------------------------------------
// Exact electrostatic potential: src/electrostatic_potential.cpp:17-361.

struct Triple {
    int t, u, v;
};

std::vector<Triple> triples(int L) {
    std::vector<Triple> out;
    for (int s = 0; s <= L; ++s)
        for (int t = s; t >= 0; --t)
            for (int u = s - t; u >= 0; --u)
                out.push_back({t, u, s - t - u});
    return out;
}

// E^{ij}_t Hermite expansion table (src/electrostatic_potential.cpp:57-93).
struct Hermite {
    int lb, width;
    std::vector<double> v;
    Hermite(int la, int lb_, double p, double pa, double pb) : lb(lb_), width(la + lb_ + 1) {
        v.assign(static_cast<std::size_t>((la + 1) * (lb + 1) * width), 0.0);
        const double half = 0.5 / p;
        at(0, 0, 0) = 1.0;
        const auto step = [&](int si, int sj, int ti, int tj, double shift) {
            for (int t = 0; t <= ti + tj; ++t) {
                double value = shift * get(si, sj, t);
                if (t > 0) value += half * get(si, sj, t - 1);
                if (t + 1 <= si + sj) value += (t + 1) * get(si, sj, t + 1);
                at(ti, tj, t) = value;
            }
        };
        for (int i = 0; i < la; ++i)
            step(i, 0, i + 1, 0, pa);
        for (int i = 0; i <= la; ++i)
            for (int j = 0; j < lb; ++j)
                step(i, j, i, j + 1, pb);
    }
    double get(int i, int j, int t) const {
        if (t < 0 || t > i + j) return 0.0;
        return v[static_cast<std::size_t>((i * (lb + 1) + j) * width + t)];
    }
    double& at(int i, int j, int t) { return v[static_cast<std::size_t>((i * (lb + 1) + j) * width + t)]; }
};

struct EspPair {
    double p;
    double Q[3];
    int L;
    std::size_t first;
};

struct Esp {
    std::vector<EspPair> pairs;
    std::vector<double> coef;
    int max_L = 0;
    std::vector<Atom> atoms;
};

std::unique_ptr<Esp> make_esp(const Molecule& M) {
    auto esp = std::make_unique<Esp>();
    esp->atoms = M.atoms;
    const auto& blocks = M.basis->blocks;
    for (const auto& b : blocks) {
        if (b.l > 13) throw std::invalid_argument("the electrostatic potential supports shells up to l = 13");
        esp->max_L = std::max(esp->max_L, 2 * b.l);
    }
    const int n = M.n;
    for (std::size_t ia = 0; ia < blocks.size(); ++ia) {
        const auto& A = blocks[ia];
        for (std::size_t ib = ia; ib < blocks.size(); ++ib) {
            const auto& B = blocks[ib];
            const int na = A.ncart(), nb = B.ncart();
            // Cartesian density block: TA^T * P(A, B) * TB (x2 off the diagonal).
            std::vector<double> pblk(static_cast<std::size_t>(A.count * B.count));
            for (int r = 0; r < A.count; ++r)
                for (int c = 0; c < B.count; ++c)
                    pblk[static_cast<std::size_t>(r * B.count + c)] = M.P[static_cast<std::size_t>(A.offset + r) * n + static_cast<std::size_t>(B.offset + c)];
            std::vector<double> left(static_cast<std::size_t>(na * B.count));
            for (int i = 0; i < na; ++i)
                for (int c = 0; c < B.count; ++c) {
                    if (A.pure.empty()) {
                        left[static_cast<std::size_t>(i * B.count + c)] = pblk[static_cast<std::size_t>(i * B.count + c)];
                    } else {
                        double acc = 0.0;
                        for (int r = 0; r < A.count; ++r)
                            acc += A.pure[static_cast<std::size_t>(r * na + i)] * pblk[static_cast<std::size_t>(r * B.count + c)];
                        left[static_cast<std::size_t>(i * B.count + c)] = acc;
                    }
                }
            std::vector<double> cart(static_cast<std::size_t>(na * nb));
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    double v;
                    if (B.pure.empty()) {
                        v = left[static_cast<std::size_t>(i * B.count + j)];
                    } else {
                        v = 0.0;
                        for (int c = 0; c < B.count; ++c)
                            v += left[static_cast<std::size_t>(i * B.count + c)] * B.pure[static_cast<std::size_t>(c * nb + j)];
                    }
                    if (ia != ib) v *= 2.0;
                    cart[static_cast<std::size_t>(i * nb + j)] = v;
                }
            if (std::all_of(cart.begin(), cart.end(), [](double v) { return v == 0.0; })) continue;
            const int la = A.l, lb = B.l, L = la + lb;
            const auto list = triples(L);
            const double AB[3] = {A.origin[0] - B.origin[0], A.origin[1] - B.origin[1], A.origin[2] - B.origin[2]};
            const double ab2 = sqnorm(AB[0], AB[1], AB[2]);
            std::vector<double> pc(list.size());
            for (int p = 0; p < A.nprim(); ++p)
                for (int q = 0; q < B.nprim(); ++q) {
                    const double al = A.exps[static_cast<std::size_t>(p)], be = B.exps[static_cast<std::size_t>(q)];
                    const double ex = al + be;
                    double Q[3];
                    for (int a = 0; a < 3; ++a)
                        Q[a] = (al * A.origin[a] + be * B.origin[a]) / ex;
                    const double scale = std::exp(-al * be / ex * ab2) * 2.0 * std::numbers::pi / ex;
                    const Hermite hx(la, lb, ex, Q[0] - A.origin[0], Q[0] - B.origin[0]);
                    const Hermite hy(la, lb, ex, Q[1] - A.origin[1], Q[1] - B.origin[1]);
                    const Hermite hz(la, lb, ex, Q[2] - A.origin[2], Q[2] - B.origin[2]);
                    std::fill(pc.begin(), pc.end(), 0.0);
                    for (int i = 0; i < na; ++i) {
                        const auto& pi = A.powers[static_cast<std::size_t>(i)];
                        for (int j = 0; j < nb; ++j) {
                            const auto& pj = B.powers[static_cast<std::size_t>(j)];
                            const double wgt = cart[static_cast<std::size_t>(i * nb + j)] * A.w(i, p) * B.w(j, q);
                            if (wgt == 0.0) continue;
                            for (std::size_t k = 0; k < list.size(); ++k) {
                                const auto& h = list[k];
                                if (h.t > pi.x + pj.x || h.u > pi.y + pj.y || h.v > pi.z + pj.z) continue;
                                pc[k] += wgt * hx.get(pi.x, pj.x, h.t) * hy.get(pi.y, pj.y, h.u) * hz.get(pi.z, pj.z, h.v);
                            }
                        }
                    }
                    if (std::all_of(pc.begin(), pc.end(), [](double c) { return c == 0.0; })) continue;
                    esp->pairs.push_back({ex, {Q[0], Q[1], Q[2]}, L, esp->coef.size()});
                    for (double c : pc)
                        esp->coef.push_back(scale * c);
                }
        }
    }
    return esp;
}

double esp_nuclear(const Esp& e, const double* r) {
    double sum = 0.0;
    for (const auto& a : e.atoms) {
        const double d = std::sqrt(sqnorm(a.pos[0] - r[0], a.pos[1] - r[1], a.pos[2] - r[2]));
        if (a.charge == 0.0 || d < 1e-8) continue;
        sum += a.charge / d;
    }
    return sum;
}

std::size_t hcount(int s) { return static_cast<std::size_t>((s + 1) * (s + 2) * (s + 3) / 6); }

// Work arrays of the Hermite recursion: two cubic tables of `side` entries
// per index, the seeds R^n_000 and the (t, u, v) in order of t + u + v.
struct HermiteWork {
    int side = 0;
    std::vector<double> prev, cur, seeds;
    std::vector<Triple> all;
    void prepare(int side_, int top) {
        if (side == side_ && all.size() >= hcount(top)) return;
        side = side_;
        const std::size_t cube = static_cast<std::size_t>(side * side * side);
        prev.assign(cube, 0.0);
        cur.assign(cube, 0.0);
        seeds.assign(static_cast<std::size_t>(top + 1), 0.0);
        all = triples(top);
    }
    std::size_t off(int t, int u, int v) const { return static_cast<std::size_t>((t * side + u) * side + v); }
};

// hermite_recursion (src/electrostatic_potential.cpp:293-324): R^0_tuv for
// t + u + v <= L into w.cur at w.off(t, u, v), from w.seeds[n] = R^n_000:
//   R^n_(t+1,u,v) = t R^(n+1)_(t-1,u,v) + X R^(n+1)_(t,u,v)
// (first nonzero index of t, u, v), from n = L down to 0.
void hermite_recursion(const double pc[3], int L, HermiteWork& w) {
    for (int n = L; n >= 0; --n) {
        std::swap(w.prev, w.cur);
        const auto& prev = w.prev;
        w.cur[0] = w.seeds[static_cast<std::size_t>(n)];
        const std::size_t entries = hcount(L - n);
        for (std::size_t k = 1; k < entries; ++k) {
            const auto& h = w.all[k];
            double value;
            if (h.t > 0) {
                value = pc[0] * prev[w.off(h.t - 1, h.u, h.v)];
                if (h.t > 1) value += (h.t - 1) * prev[w.off(h.t - 2, h.u, h.v)];
            } else if (h.u > 0) {
                value = pc[1] * prev[w.off(h.t, h.u - 1, h.v)];
                if (h.u > 1) value += (h.u - 1) * prev[w.off(h.t, h.u - 2, h.v)];
            } else {
                value = pc[2] * prev[w.off(h.t, h.u, h.v - 1)];
                if (h.v > 1) value += (h.v - 1) * prev[w.off(h.t, h.u, h.v - 2)];
            }
            w.cur[w.off(h.t, h.u, h.v)] = value;
        }
    }
}

// hermite_integrals (src/electrostatic_potential.cpp:269-280): seeds
// R^n_000 = (-2p)^n F_n(p |PC|^2), then the recursion to order L.
void hermite_integrals(double p, const double pc[3], int L, HermiteWork& w) {
    boys(p * sqnorm(pc[0], pc[1], pc[2]), L, w.seeds.data());
    double power = 1.0;
    for (int n = 1; n <= L; ++n) {
        power *= -2.0 * p;
        w.seeds[static_cast<std::size_t>(n)] *= power;
    }
    hermite_recursion(pc, L, w);
}

// pair_potential (src/electrostatic_potential.cpp:326-338): sum_k c_k R^0_k
// of one pair at the point r.
double pair_potential(const Esp& e, const EspPair& pr, const double* r, HermiteWork& w) {
    const double pc[3] = {pr.Q[0] - r[0], pr.Q[1] - r[1], pr.Q[2] - r[2]};
    hermite_integrals(pr.p, pc, pr.L, w);
    const double* c = e.coef.data() + pr.first;
    const std::size_t nk = hcount(pr.L);
    double sum = 0.0;
    for (std::size_t k = 0; k < nk; ++k)
        sum += c[k] * w.cur[w.off(w.all[k].t, w.all[k].u, w.all[k].v)];
    return sum;
}

// electronic (src/electrostatic_potential.cpp:340-348): minus the sum over
// every pair.
double esp_electronic(const Esp& e, const double* r) {
    HermiteWork w;
    w.prepare(e.max_L + 1, e.max_L);
    double total = 0.0;
    for (const auto& pr : e.pairs)
        total += pair_potential(e, pr, r, w);
    return -total;
}

// ---------------------------------------------------------------------------
// Far-field approximation of the electrostatic potential: error bounds and
// plans (src/electrostatic_potential.cpp:363-699), cluster evaluation
// (:701-844) and target clusters (src/potential_clusters.hpp:29-105,
// src/evaluator.cpp:391-479).

constexpr int kLocalOrder = 6;      // src/electrostatic_potential.hpp:64
constexpr int kMultipoleOrder = 10; // src/electrostatic_potential.hpp:66
constexpr int kLevelCount = 33;     // src/electrostatic_potential.hpp:76
constexpr double kSourceCell = 2.0; // src/electrostatic_potential.hpp:77
constexpr double kPointCube = 1.6;  // include/chem/kumo/evaluator.hpp:170
constexpr int kTileSide = 8;        // src/evaluator.cpp:399

// Factor from a stored Hermite coefficient (2 pi / p included) to the weight
// of its point multipole, sqrt(pi / p) / 2 (src/electrostatic_potential.cpp:31-33).
double to_multipole(double p) { return 0.5 * std::sqrt(std::numbers::pi / p); }

// G_m(r), m < count: bound on every order-m derivative of erfc(sqrt(p) R) / R
// for R >= r (src/electrostatic_potential.cpp:385-430), through the upper
// incomplete gamma function Gamma((m + 1) / 2, p r^2 / 2), advanced from
// Gamma(1/2, x) = sqrt(pi) erfc(sqrt(x)) and Gamma(1, x) = exp(-x) by
// Gamma(s + 1, x) = s Gamma(s, x) + x^s exp(-x).
void erfc_bounds(double p, double r, std::size_t count, double* out) {
    if (!(r > 0.0)) {
        std::fill(out, out + count, std::numeric_limits<double>::infinity());
        return;
    }
    const double x = 0.5 * p * r * r;
    const double decay = std::exp(-x), rx = std::sqrt(x);
    const double scale = 2.0 / (r * r), rscale = std::sqrt(scale);
    double g_even = std::sqrt(std::numbers::pi) * std::erfc(rx), x_even = rx, s_even = 0.5;
    double g_odd = decay, x_odd = x, s_odd = 1.0;
    double spow = rscale, norm2 = 1.0;
    constexpr double k = 1.0865; // Cramer's bound on Hermite functions
    const double pre = 2.0 / std::sqrt(std::numbers::pi) * k * k * k * 0.5;
    for (std::size_t m = 0; m < count; ++m) {
        if (m > 0) {
            norm2 *= 2.0 * static_cast<double>(m);
            spow *= rscale;
        }
        double g;
        if (m % 2 == 0) {
            if (m > 0) {
                g_even = s_even * g_even + x_even * decay;
                x_even *= x;
                s_even += 1.0;
            }
            g = g_even;
        } else {
            if (m > 1) {
                g_odd = s_odd * g_odd + x_odd * decay;
                x_odd *= x;
                s_odd += 1.0;
            }
            g = g_odd;
        }
        out[m] = pre * std::sqrt(norm2) * spow * g;
    }
}

// T_L(d) = sum_(n > K) (n + L)! / n! rho^n / d^(n + L + 1), the Taylor
// remainder bound of the order-L multipole terms (src/electrostatic_potential.cpp:432-455):
// infinite unless d > 1.25 rho; summed until the term ratio drops below 0.9,
// then closed by the geometric tail.
double taylor_tail(int L, int K, double rho, double d) {
    if (!(d > 1.25 * rho)) return std::numeric_limits<double>::infinity();
    const double q = rho / d;
    double term = 1.0 / std::pow(d, L + 1);
    for (int j = 1; j <= L; ++j)
        term *= j;
    for (int n = 1; n <= K + 1; ++n)
        term *= (n + L) / static_cast<double>(n) * q;
    double sum = 0.0;
    for (int n = K + 1; n < K + 1000; ++n) {
        sum += term;
        const double ratio = (n + L + 1) / static_cast<double>(n + 1) * q;
        if (ratio < 0.9) return sum + term * ratio / (1.0 - ratio);
        term *= ratio;
        if (!std::isfinite(term)) break;
    }
    return std::numeric_limits<double>::infinity();
}

// Smallest d in [low, high] with bound(d) <= allowed for a decreasing bound
// (src/electrostatic_potential.cpp:457-472): `high` doubled up to 1e4 bohr,
// then bisection to a relative 1e-3 (at most 40 steps), rounded up.
template <typename Bound> double smallest_distance(const Bound& bound, double low, double high, double allowed) {
    while (high < 1e4 && !(bound(high) <= allowed))
        high *= 2.0;
    if (!(bound(high) <= allowed)) return std::numeric_limits<double>::infinity();
    for (int it = 0; it < 40 && high - low > 1e-3 * high; ++it) {
        const double mid = 0.5 * (low + high);
        if (bound(mid) <= allowed)
            high = mid;
        else
            low = mid;
    }
    return high;
}

// A double rounded up to a float (src/electrostatic_potential.cpp:474-481).
float round_up_float(double v) {
    const float f = static_cast<float>(v);
    return static_cast<double>(f) < v ? std::nextafter(f, std::numeric_limits<float>::infinity()) : f;
}

// W_L: sum of |e_k| over the Hermite terms of order L of a pair
// (multipole_weights, src/electrostatic_potential.cpp:484-495).
void multipole_weights(const Esp& e, const EspPair& pr, std::vector<double>& w) {
    w.assign(static_cast<std::size_t>(pr.L + 1), 0.0);
    const double* c = e.coef.data() + pr.first;
    const double f = to_multipole(pr.p);
    std::size_t k = 0;
    for (int order = 0; order <= pr.L; ++order)
        for (int j = 0; j < (order + 1) * (order + 2) / 2; ++j, ++k)
            w[static_cast<std::size_t>(order)] += std::abs(c[k]) * f;
}

struct SourceCluster {
    double c[3] = {0.0, 0.0, 0.0};
    double radius = 0.0, min_distance = 0.0;
    std::vector<std::size_t> pairs;
    std::vector<double> moments; // (N + 1) blocks of moment_count: moments truncated at |delta| <= m
};

struct FarPlan {
    std::size_t npairs = 0;
    double radius = 0.0, tolerance = 0.0, pair_tolerance = 0.0;
    double levels[kLevelCount] = {};
    std::vector<float> level_d2; // npairs x kLevelCount
    std::vector<SourceCluster> clusters;
    std::size_t moment_count = 0;
};

// plan_pair_levels (src/electrostatic_potential.cpp:503-559): the error
// bound of each pair's local expansion (order K about a cluster of radius
// rho) at 64 distances from max(1.25 rho, 0.05) to 400 bohr in geometric
// progression,
//   sum_L W_L [T_L(d) + G_L(d - rho) + sum_(n<=K) (sqrt(3) rho)^n / n! G_(L+n)(d)],
// and for each level the first of these distances whose bound is within it.
void plan_pair_levels(const Esp& e, FarPlan& plan) {
    constexpr int K = kLocalOrder, G = 64;
    const double rho = plan.radius;
    const double first = std::max(1.25 * rho, 0.05);
    const double ratio = std::pow(400.0 / first, 1.0 / (G - 1));
    double dist[G], bound[G];
    for (int g = 0; g < G; ++g)
        dist[g] = first * std::pow(ratio, g);
    std::vector<double> w, nearb, farb;
    for (std::size_t i = 0; i < e.pairs.size(); ++i) {
        const EspPair& pr = e.pairs[i];
        multipole_weights(e, pr, w);
        nearb.assign(static_cast<std::size_t>(pr.L + 1), 0.0);
        farb.assign(static_cast<std::size_t>(pr.L + K + 1), 0.0);
        for (int g = 0; g < G; ++g) {
            const double d = dist[g];
            erfc_bounds(pr.p, d - rho, nearb.size(), nearb.data());
            erfc_bounds(pr.p, d, farb.size(), farb.data());
            double total = 0.0;
            for (int L = 0; L <= pr.L; ++L) {
                const double wl = w[static_cast<std::size_t>(L)];
                if (wl == 0.0) continue;
                double term = taylor_tail(L, K, rho, d) + nearb[static_cast<std::size_t>(L)];
                double sc = 1.0;
                for (int n = 0; n <= K; ++n) {
                    if (n > 0) sc *= std::sqrt(3.0) * rho / n;
                    term += sc * farb[static_cast<std::size_t>(L + n)];
                }
                total += wl * term;
            }
            bound[g] = total;
        }
        int g = 0;
        for (int j = 0; j < kLevelCount; ++j) {
            while (g < G && !(bound[g] <= plan.levels[j]))
                ++g;
            if (g == G) break;
            plan.level_d2[i * kLevelCount + static_cast<std::size_t>(j)] = round_up_float(dist[g] * dist[g]);
        }
    }
}

// plan_source_clusters (src/electrostatic_potential.cpp:561-699): pairs
// grouped by the cube of edge 2 bohr holding their centre and by exponent
// interval [4^k, 4^(k+1)), clusters numbered by first appearance; each with
// the centre of the bounding box of its pair centres, the largest distance
// of a pair centre from it, the moments
//   Q^(m)_nu = sum over pairs and terms alpha_k + delta = nu, |delta| <= m,
//              of e_k u^delta / delta!   (u = pair centre - cluster centre)
// for m = 0..N, and the distance beyond which the translation to a target
// cluster is within a quarter of the tolerance shared among the clusters.
void plan_source_clusters(const Esp& e, FarPlan& plan) {
    constexpr int N = kMultipoleOrder;
    std::map<std::array<long, 4>, std::size_t> index;
    for (std::size_t i = 0; i < e.pairs.size(); ++i) {
        const EspPair& pr = e.pairs[i];
        std::array<long, 4> key{};
        for (int a = 0; a < 3; ++a)
            key[static_cast<std::size_t>(a)] = static_cast<long>(std::floor(pr.Q[a] / kSourceCell));
        key[3] = static_cast<long>(std::floor(std::log2(pr.p) / 2.0));
        const auto it = index.find(key);
        std::size_t c;
        if (it == index.end()) {
            c = plan.clusters.size();
            index.emplace(key, c);
            plan.clusters.emplace_back();
        } else {
            c = it->second;
        }
        plan.clusters[c].pairs.push_back(i);
    }
    const double share = 0.25 * plan.tolerance / static_cast<double>(plan.clusters.size());
    const int side = e.max_L + N + 1;
    const auto nu_list = triples(e.max_L + N);
    plan.moment_count = nu_list.size();
    std::vector<int> compact(static_cast<std::size_t>(side * side * side), -1);
    for (std::size_t m = 0; m < nu_list.size(); ++m)
        compact[static_cast<std::size_t>((nu_list[m].t * side + nu_list[m].u) * side + nu_list[m].v)] = static_cast<int>(m);
    const auto deltas = triples(N);
    std::vector<double> w, cw, g;
    for (auto& cl : plan.clusters) {
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a)
            lo[a] = hi[a] = e.pairs[cl.pairs.front()].Q[a];
        double pmin = std::numeric_limits<double>::infinity();
        int top = 0;
        for (const std::size_t i : cl.pairs) {
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], e.pairs[i].Q[a]);
                hi[a] = std::max(hi[a], e.pairs[i].Q[a]);
            }
            pmin = std::min(pmin, e.pairs[i].p);
            top = std::max(top, e.pairs[i].L);
        }
        for (int a = 0; a < 3; ++a)
            cl.c[a] = 0.5 * (lo[a] + hi[a]);
        cl.radius = 0.0;
        cw.assign(static_cast<std::size_t>(top + 1), 0.0);
        std::vector<double> by_order(static_cast<std::size_t>(N + 1) * plan.moment_count, 0.0);
        for (const std::size_t i : cl.pairs) {
            const EspPair& pr = e.pairs[i];
            const double u[3] = {pr.Q[0] - cl.c[0], pr.Q[1] - cl.c[1], pr.Q[2] - cl.c[2]};
            cl.radius = std::max(cl.radius, std::sqrt(sqnorm(u[0], u[1], u[2])));
            multipole_weights(e, pr, w);
            for (std::size_t o = 0; o < w.size(); ++o)
                cw[o] += w[o];
            const double f = to_multipole(pr.p);
            const double* c = e.coef.data() + pr.first;
            const std::size_t nk = hcount(pr.L);
            for (std::size_t k = 0; k < nk; ++k) {
                const double ek = c[k] * f;
                if (ek == 0.0) continue;
                const Triple& al = nu_list[k];
                for (const Triple& d : deltas) {
                    double term = ek;
                    for (int j = 1; j <= d.t; ++j)
                        term *= u[0] / j;
                    for (int j = 1; j <= d.u; ++j)
                        term *= u[1] / j;
                    for (int j = 1; j <= d.v; ++j)
                        term *= u[2] / j;
                    const int nu = compact[static_cast<std::size_t>(((al.t + d.t) * side + al.u + d.u) * side + al.v + d.v)];
                    by_order[static_cast<std::size_t>(d.t + d.u + d.v) * plan.moment_count + static_cast<std::size_t>(nu)] += term;
                }
            }
        }
        cl.moments.assign(by_order.size(), 0.0);
        for (int m = 0; m <= N; ++m)
            for (std::size_t nu = 0; nu < plan.moment_count; ++nu) {
                const std::size_t at = static_cast<std::size_t>(m) * plan.moment_count + nu;
                cl.moments[at] = (m > 0 ? cl.moments[at - plan.moment_count] : 0.0) + by_order[at];
            }
        const double reach = cl.radius + plan.radius;
        const auto truncation = [&](double D) {
            double total = 0.0;
            for (std::size_t o = 0; o < cw.size(); ++o)
                if (cw[o] > 0.0) total += cw[o] * taylor_tail(static_cast<int>(o), N, reach, D);
            return total;
        };
        g.assign(cw.size(), 0.0);
        const auto erfc_terms = [&](double r) {
            erfc_bounds(pmin, r, g.size(), g.data());
            double total = 0.0;
            for (std::size_t o = 0; o < cw.size(); ++o)
                total += cw[o] * g[o];
            return total;
        };
        const double dt = smallest_distance(truncation, 1.25 * reach, 2.5 * reach + 1e-3, share);
        const double re = smallest_distance(erfc_terms, 0.0, 1.0 / std::sqrt(pmin), share);
        cl.min_distance = std::max(dt, reach + re);
    }
}

// far_field_plan (src/electrostatic_potential.cpp:497-501 and its body
// above :503): levels pair_tolerance * 10^(-j/2) with pair_tolerance half
// the tolerance; tolerance 0 plans nothing (every pair exact).
std::unique_ptr<FarPlan> far_field_plan(const Esp& e, double radius, double tolerance) {
    if (!std::isfinite(radius) || radius < 0.0) throw std::invalid_argument("cluster radius must be finite and nonnegative");
    if (!std::isfinite(tolerance) || tolerance < 0.0)
        throw std::invalid_argument("electrostatic potential tolerance must be finite and nonnegative");
    auto plan = std::make_unique<FarPlan>();
    plan->npairs = e.pairs.size();
    plan->radius = radius;
    plan->tolerance = tolerance;
    plan->level_d2.assign(e.pairs.size() * kLevelCount, std::numeric_limits<float>::infinity());
    if (tolerance == 0.0 || e.pairs.empty()) return plan;
    plan->pair_tolerance = 0.5 * tolerance;
    for (int j = 0; j < kLevelCount; ++j)
        plan->levels[j] = plan->pair_tolerance * std::pow(10.0, -0.5 * j);
    plan_pair_levels(e, *plan);
    plan_source_clusters(e, *plan);
    return plan;
}

// cluster_plan (src/potential_clusters.hpp:29-35): the radius enlarged by a
// relative 1e-10 plus 1e-12 bohr for rounding in the coordinates.
std::unique_ptr<FarPlan> cluster_plan(const Esp& e, double radius, double tolerance) {
    return far_field_plan(e, radius * (1.0 + 1e-10) + 1e-12, tolerance);
}

struct FarStats {
    std::int64_t far_clusters = 0, expanded = 0, near = 0;
};

// expand_far_field (src/electrostatic_potential.cpp:730-826): the local
// expansion about `b` (factors (-1)^|beta| / beta! applied) in `local`, from
// the source clusters at least min_distance away (one multipole-to-local
// translation each, local_beta += sum_nu Q^(N - |beta|)_nu d^(nu + beta) 1/|X|)
// and from the remaining pairs whose error levels fit the pair budget
// (smallest levels first; at the level that fits only in part, pairs in index
// order); returns the pairs to evaluate exactly.
std::vector<std::size_t> expand_far_field(const Esp& e, const FarPlan& plan, const double b[3], HermiteWork& w,
                                          std::vector<double>& local, FarStats& st) {
    constexpr int N = kMultipoleOrder;
    const int top = e.max_L + N;
    w.prepare(top + 1, top);
    const std::size_t nlocal = hcount(N);
    local.assign(nlocal, 0.0);
    std::vector<char> handled(e.pairs.size(), 0);
    for (const auto& cl : plan.clusters) {
        const double x[3] = {cl.c[0] - b[0], cl.c[1] - b[1], cl.c[2] - b[2]};
        if (!(std::sqrt(sqnorm(x[0], x[1], x[2])) >= cl.min_distance)) continue;
        ++st.far_clusters;
        // coulomb_derivatives (src/electrostatic_potential.cpp:282-291):
        // seeds (-1)^n (2n - 1)!! / |X|^(2n + 1).
        const double inv = 1.0 / std::sqrt(sqnorm(x[0], x[1], x[2])), inv2 = inv * inv;
        double seed = inv;
        for (int n = 0; n <= top; ++n) {
            w.seeds[static_cast<std::size_t>(n)] = seed;
            seed *= -(2.0 * n + 1.0) * inv2;
        }
        hermite_recursion(x, top, w);
        for (std::size_t bi = 0; bi < nlocal; ++bi) {
            const Triple& be = w.all[bi];
            const int m = N - (be.t + be.u + be.v);
            const double* q = cl.moments.data() + static_cast<std::size_t>(m) * plan.moment_count;
            const std::size_t terms = hcount(e.max_L + m);
            double sum = 0.0;
            for (std::size_t nu = 0; nu < terms; ++nu) {
                const Triple& a = w.all[nu];
                sum += q[nu] * w.cur[w.off(a.t + be.t, a.u + be.u, a.v + be.v)];
            }
            local[bi] += sum;
        }
        for (const std::size_t i : cl.pairs)
            handled[i] = 1;
    }
    std::vector<signed char> level(e.pairs.size(), -1);
    std::array<std::size_t, kLevelCount> population{};
    for (std::size_t i = 0; i < e.pairs.size(); ++i) {
        if (handled[i]) continue;
        const EspPair& pr = e.pairs[i];
        const double d2 = sqnorm(pr.Q[0] - b[0], pr.Q[1] - b[1], pr.Q[2] - b[2]);
        const float* lim = plan.level_d2.data() + i * kLevelCount;
        int j = -1;
        while (j + 1 < kLevelCount && d2 >= static_cast<double>(lim[j + 1]))
            ++j;
        level[i] = static_cast<signed char>(j);
        if (j >= 0) ++population[static_cast<std::size_t>(j)];
    }
    int cut = kLevelCount;
    std::size_t partial = 0;
    double used = 0.0;
    for (int j = kLevelCount - 1; j >= 0; --j) {
        const double v = plan.levels[j];
        const std::size_t n = population[static_cast<std::size_t>(j)];
        if (used + static_cast<double>(n) * v <= plan.pair_tolerance) {
            used += static_cast<double>(n) * v;
            cut = j - 1;
            continue;
        }
        partial = static_cast<std::size_t>(std::floor((plan.pair_tolerance - used) / v));
        cut = j;
        break;
    }
    std::vector<std::size_t> near;
    const std::size_t nlow = hcount(kLocalOrder);
    for (std::size_t i = 0; i < e.pairs.size(); ++i) {
        if (handled[i]) continue;
        bool expand = level[i] >= 0 && level[i] > cut;
        if (!expand && level[i] >= 0 && level[i] == cut && partial > 0) {
            expand = true;
            --partial;
        }
        if (!expand) {
            near.push_back(i);
            continue;
        }
        ++st.expanded;
        const EspPair& pr = e.pairs[i];
        const double pc[3] = {pr.Q[0] - b[0], pr.Q[1] - b[1], pr.Q[2] - b[2]};
        hermite_integrals(pr.p, pc, pr.L + kLocalOrder, w);
        const double* c = e.coef.data() + pr.first;
        const std::size_t nk = hcount(pr.L);
        for (std::size_t k = 0; k < nk; ++k) {
            const Triple& a = w.all[k];
            for (std::size_t bi = 0; bi < nlow; ++bi) {
                const Triple& be = w.all[bi];
                local[bi] += c[k] * w.cur[w.off(a.t + be.t, a.u + be.u, a.v + be.v)];
            }
        }
    }
    for (std::size_t bi = 0; bi < nlocal; ++bi) {
        const Triple& be = w.all[bi];
        double f = ((be.t + be.u + be.v) % 2 == 0) ? 1.0 : -1.0;
        for (int j = 2; j <= be.t; ++j)
            f /= j;
        for (int j = 2; j <= be.u; ++j)
            f /= j;
        for (int j = 2; j <= be.v; ++j)
            f /= j;
        local[bi] *= f;
    }
    st.near += static_cast<std::int64_t>(near.size());
    return near;
}

// local_polynomial (src/electrostatic_potential.cpp:828-844): the local
// expansion at displacement h from its centre.
double local_polynomial(const std::vector<double>& local, const std::vector<Triple>& all, const double h[3]) {
    double pw[3][kMultipoleOrder + 1];
    for (int a = 0; a < 3; ++a) {
        pw[a][0] = 1.0;
        for (int k = 1; k <= kMultipoleOrder; ++k)
            pw[a][k] = pw[a][k - 1] * h[a];
    }
    double poly = 0.0;
    for (std::size_t bi = 0; bi < local.size(); ++bi) {
        const Triple& be = all[bi];
        poly += local[bi] * pw[0][be.t] * pw[1][be.u] * pw[2][be.v];
    }
    return poly;
}

// evaluate_cluster (src/electrostatic_potential.cpp:701-728): points within
// the plan's radius of the centre (with rounding slack), each the nuclear
// term minus (local polynomial + the exact near pairs).
void evaluate_cluster(const Esp& e, const FarPlan& plan, std::int64_t n, const double* pts, const double centre[3],
                      double* out, FarStats& st) {
    if (plan.level_d2.size() != e.pairs.size() * kLevelCount)
        throw std::invalid_argument("far-field plan does not match the potential");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(centre[a])) throw std::invalid_argument("cluster centre must be finite");
    const double cn = std::sqrt(sqnorm(centre[0], centre[1], centre[2]));
    for (std::int64_t j = 0; j < n; ++j) {
        const double* r = pts + 3 * j;
        if (!std::isfinite(r[0]) || !std::isfinite(r[1]) || !std::isfinite(r[2]))
            throw std::invalid_argument("electrostatic potential point must be finite");
        const double d = std::sqrt(sqnorm(r[0] - centre[0], r[1] - centre[1], r[2] - centre[2]));
        if (d > plan.radius + 1e-12 * (plan.radius + cn)) throw std::invalid_argument("cluster point outside the planned radius");
    }
    HermiteWork w;
    std::vector<double> local;
    const auto near = expand_far_field(e, plan, centre, w, local, st);
    for (std::int64_t j = 0; j < n; ++j) {
        const double* r = pts + 3 * j;
        const double h[3] = {r[0] - centre[0], r[1] - centre[1], r[2] - centre[2]};
        double sum = local_polynomial(local, w.all, h);
        for (const std::size_t i : near)
            sum += pair_potential(e, e.pairs[i], r, w);
        out[j] = esp_nuclear(e, r) - sum;
    }
}

// ---------------------------------------------------------------------------
// Points and the evaluation driver: src/grid_geometry.hpp, src/evaluator.cpp.

struct PointSet {
    const double* origin = nullptr;
    const double* steps = nullptr; // column-major 3x3
    const int* counts = nullptr;
    const double* pts = nullptr;
    std::int64_t n = 0;
    bool grid() const { return pts == nullptr; }
    // grid_point (src/grid_geometry.hpp:59-68): origin + steps * (i, j, k).
    void point(std::int64_t q, double* out) const {
        if (!grid()) {
            out[0] = pts[3 * q];
            out[1] = pts[3 * q + 1];
            out[2] = pts[3 * q + 2];
            return;
        }
        const std::int64_t ny = counts[1], nz = counts[2];
        const double f[3] = {static_cast<double>(q / (ny * nz)), static_cast<double>((q / nz) % ny),
                             static_cast<double>(q % nz)};
        for (int a = 0; a < 3; ++a)
            out[a] = origin[a] + ((steps[a] * f[0] + steps[3 + a] * f[1]) + steps[6 + a] * f[2]);
    }
    void block(std::int64_t start, int size, std::vector<double>& out) const {
        out.resize(static_cast<std::size_t>(3 * size));
        for (int j = 0; j < size; ++j)
            point(start + j, out.data() + 3 * j);
    }
};

bool uses_point_path(const ComponentPlan& c, int backend, bool grid) {
    return c.derivatives() || ((backend == AO_MO || !grid) && c.density);
}
bool on_grid_path(const orc_property& p, int backend, bool grid) {
    return grid && (p.kind == Orbital || (p.kind == Density && backend == AO_P));
}
bool from_point_path(const orc_property& p, int backend, bool grid) {
    return !on_grid_path(p, backend, grid) && !is_orbital_property(p.kind) && p.kind != ElectrostaticPotential &&
           !is_orbital_energy(p.kind);
}

// ---------------------------------------------------------------------------
// Orbital-energy fields: OrbitalEnergyFields (src/orbital_energy_fields.hpp:66-262)
// and store_orbital_energy_block (src/evaluator.cpp:481-520).

// One weighted set of MO columns: each column a coefficient vector (already
// scaled by the square root of its occupation or weight) with an energy.
struct OrbitalColumns {
    std::vector<std::vector<double>> col;
    std::vector<double> energy;
    void add(const std::vector<double>& C, int nmo, int mo, int n, double scale, double e) {
        std::vector<double> c(static_cast<std::size_t>(n));
        for (int a = 0; a < n; ++a)
            c[static_cast<std::size_t>(a)] = scale * C[static_cast<std::size_t>(a) * nmo + mo];
        col.push_back(std::move(c));
        energy.push_back(e);
    }
};

// Normalized Gaussian weights of the orbital-weighted Fukui functions
// (orbital_weights, src/orbital_energy_fields.hpp:75-96): exponents shifted
// by their largest value before exponentiation.
std::vector<double> fukui_weights(const std::vector<double>& e, double mu, double delta) {
    std::vector<double> x(e.size()), w(e.size());
    for (std::size_t i = 0; i < e.size(); ++i) {
        const double t = (mu - e[i]) / delta;
        x[i] = -t * t;
    }
    const double top = *std::max_element(x.begin(), x.end());
    double sum = 0.0;
    for (std::size_t i = 0; i < e.size(); ++i) {
        w[i] = std::exp(x[i] - top);
        sum += w[i];
    }
    for (double& v : w)
        v /= sum;
    return w;
}

struct OrbitalEnergySet {
    bool virtuals = false, fukui = false;
    OrbitalColumns occupied, virt, attachment, plus, minus;
};

OrbitalEnergySet orbital_energy_set(const Molecule& M, const Plan& plan, double delta) {
    const std::string what = plan.orbital_weighted_fukui ? "the orbital-weighted Fukui function or dual descriptor"
                             : plan.virtual_orbitals     ? "the local electron affinity and related fields"
                                                         : "the average local ionization energy";
    OrbitalEnergySet s;
    s.virtuals = plan.virtual_orbitals;
    s.fukui = plan.orbital_weighted_fukui;
    const auto bad = [&](const std::string& tail) { throw std::invalid_argument(what + tail); };
    if (s.virtuals && !M.unrestricted && M.nalpha != M.nbeta) bad(" is not defined for a restricted open-shell calculation");
    if (s.fukui && (M.unrestricted || M.nalpha != M.nbeta)) bad(" is defined only for a restricted closed-shell calculation");
    if (s.fukui && (!std::isfinite(delta) || delta <= 0.0)) bad(": the Fukui delta must be finite and positive");
    if (!M.eligible) bad(" needs occupied MOs that reproduce the FCHK density matrices (an SCF calculation)");
    const int nmo_b = M.unrestricted ? M.nmo_b : M.nmo_a;
    if (M.Ea.size() < static_cast<std::size_t>(M.nmo_a) ||
        (M.unrestricted && M.Eb.size() < static_cast<std::size_t>(nmo_b)) ||
        M.Ea.size() < static_cast<std::size_t>(M.nalpha) ||
        (M.unrestricted ? M.Eb : M.Ea).size() < static_cast<std::size_t>(M.nbeta))
        bad(" needs the FCHK orbital energies");
    const auto channel = [&](const std::vector<double>& C, int nmo, const std::vector<double>& e, int occupied,
                             int doubly, double weight) {
        for (int i = 0; i < nmo; ++i) {
            const double ei = e[static_cast<std::size_t>(i)];
            if (i < occupied) {
                s.occupied.add(C, nmo, i, M.n, std::sqrt(i < doubly ? 2.0 : 1.0), std::abs(ei));
            } else {
                s.virt.add(C, nmo, i, M.n, 1.0, ei);
                if (ei < 0.0) s.attachment.add(C, nmo, i, M.n, std::sqrt(weight), ei);
            }
        }
    };
    if (M.unrestricted) {
        channel(M.Ca, M.nmo_a, M.Ea, M.nalpha, 0, 1.0);
        channel(M.Cb, M.nmo_b, M.Eb, M.nbeta, 0, 1.0);
    } else {
        channel(M.Ca, M.nmo_a, M.Ea, std::max(M.nalpha, M.nbeta), std::min(M.nalpha, M.nbeta), 2.0);
    }
    if (s.virtuals && s.virt.col.empty()) bad(" needs virtual orbitals, and the FCHK has none");
    if (s.fukui) {
        const int count = M.nmo_a, homo = M.nalpha;
        if (homo == 0 || homo >= count) bad(" needs occupied and virtual orbitals, and the FCHK lacks one kind");
        const double mu = 0.5 * (M.Ea[static_cast<std::size_t>(homo - 1)] + M.Ea[static_cast<std::size_t>(homo)]);
        const auto columns = [&](int first, int last, OrbitalColumns& out) {
            const std::vector<double> range(M.Ea.begin() + first, M.Ea.begin() + last);
            const auto w = fukui_weights(range, mu, delta);
            for (int i = first; i < last; ++i)
                if (w[static_cast<std::size_t>(i - first)] > 0.0)
                    out.add(M.Ca, M.nmo_a, i, M.n, std::sqrt(w[static_cast<std::size_t>(i - first)]), 0.0);
        };
        columns(homo, count, s.plus);
        columns(0, homo, s.minus);
    }
    return s;
}

// Weighted square sums of one column set at the points of a block:
// density[t] = sum_i (c_i . phi_t)^2, energy[t] = sum_i e_i (c_i . phi_t)^2.
void column_sums(const OrbitalColumns& oc, const Panels& phi, int n, int size, std::vector<double>& density,
                 std::vector<double>* energy) {
    density.assign(static_cast<std::size_t>(size), 0.0);
    if (energy) energy->assign(static_cast<std::size_t>(size), 0.0);
    for (int t = 0; t < size; ++t) {
        double d = 0.0, e = 0.0;
        for (std::size_t i = 0; i < oc.col.size(); ++i) {
            double v = 0.0;
            for (int a = 0; a < n; ++a)
                v += oc.col[i][static_cast<std::size_t>(a)] * phi.at(a, t);
            const double sq = v * v;
            d += sq;
            e += oc.energy[i] * sq;
        }
        density[static_cast<std::size_t>(t)] = d;
        if (energy) (*energy)[static_cast<std::size_t>(t)] = e;
    }
}

template <typename Body> void for_blocks(std::int64_t n, int bs, const Body& body) {
    const std::int64_t nb = n / bs + (n % bs != 0);
    std::exception_ptr err;
#pragma omp parallel for schedule(dynamic) if (nb > 1)
    for (std::int64_t b = 0; b < nb; ++b) {
        try {
            const std::int64_t start = b * bs;
            body(start, static_cast<int>(std::min<std::int64_t>(bs, n - start)));
        } catch (...) {
#pragma omp critical(orc_err)
            if (!err) err = std::current_exception();
        }
    }
    if (err) std::rethrow_exception(err);
}

// The potential on a grid in tiles of up to 8 x 8 x 8 points, each a target
// cluster about the centre of its box of grid indices, with one plan whose
// radius covers the largest tile (src/evaluator.cpp:391-443,
// grid_box_radius src/potential_clusters.hpp:37-49).
void esp_grid_clusters(const Esp& e, const PointSet& set, double tolerance, double* out) {
    int tiles[3];
    double edges[3];
    for (int a = 0; a < 3; ++a) {
        edges[a] = std::min(kTileSide, set.counts[a]) - 1;
        tiles[a] = (set.counts[a] + kTileSide - 1) / kTileSide;
    }
    double radius = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        double o[3], v[3];
        for (int a = 0; a < 3; ++a)
            o[a] = ((corner >> a) & 1 ? 0.5 : -0.5) * edges[a];
        for (int a = 0; a < 3; ++a)
            v[a] = (set.steps[a] * o[0] + set.steps[3 + a] * o[1]) + set.steps[6 + a] * o[2];
        radius = std::max(radius, std::sqrt(sqnorm(v[0], v[1], v[2])));
    }
    const auto plan = cluster_plan(e, radius, tolerance);
    const std::int64_t ntiles = static_cast<std::int64_t>(tiles[0]) * tiles[1] * tiles[2];
    const std::int64_t ny = set.counts[1], nz = set.counts[2];
    for_blocks(ntiles, 1, [&](std::int64_t index, int) {
        const int t[3] = {static_cast<int>(index / (static_cast<std::int64_t>(tiles[1]) * tiles[2])),
                          static_cast<int>((index / tiles[2]) % tiles[1]), static_cast<int>(index % tiles[2])};
        int begin[3], end[3];
        for (int a = 0; a < 3; ++a) {
            begin[a] = t[a] * kTileSide;
            end[a] = std::min(begin[a] + kTileSide, set.counts[a]);
        }
        std::vector<double> pts;
        std::vector<std::int64_t> target;
        for (int i = begin[0]; i < end[0]; ++i)
            for (int j = begin[1]; j < end[1]; ++j)
                for (int k = begin[2]; k < end[2]; ++k) {
                    const std::int64_t q = (static_cast<std::int64_t>(i) * ny + j) * nz + k;
                    double r[3];
                    set.point(q, r);
                    pts.insert(pts.end(), r, r + 3);
                    target.push_back(q);
                }
        const double f[3] = {0.5 * (begin[0] + end[0] - 1), 0.5 * (begin[1] + end[1] - 1), 0.5 * (begin[2] + end[2] - 1)};
        double c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = set.origin[a] + ((set.steps[a] * f[0] + set.steps[3 + a] * f[1]) + set.steps[6 + a] * f[2]);
        std::vector<double> v(target.size());
        FarStats st;
        evaluate_cluster(e, *plan, static_cast<std::int64_t>(target.size()), pts.data(), c, v.data(), st);
        for (std::size_t q = 0; q < target.size(); ++q)
            out[target[q]] = v[q];
    });
}

// The potential at arbitrary points grouped by the axis-aligned cube of edge
// 1.6 bohr holding them, each cube a target cluster about its centre, one
// plan covering half a cube diagonal (src/evaluator.cpp:445-479, cluster_by
// src/potential_clusters.hpp:51-76).
void esp_point_clusters(const Esp& e, std::int64_t n, const double* points, double tolerance, double* out) {
    using Cell = std::array<std::int64_t, 3>;
    std::vector<Cell> cells(static_cast<std::size_t>(n));
    for (std::int64_t q = 0; q < n; ++q)
        for (int a = 0; a < 3; ++a) {
            const double index = std::floor(points[3 * q + a] / kPointCube);
            if (!(std::abs(index) < 1e15)) throw std::invalid_argument("point too far for the electrostatic potential");
            cells[static_cast<std::size_t>(q)][static_cast<std::size_t>(a)] = static_cast<std::int64_t>(index);
        }
    std::vector<std::int64_t> order(static_cast<std::size_t>(n));
    for (std::int64_t q = 0; q < n; ++q)
        order[static_cast<std::size_t>(q)] = q;
    std::stable_sort(order.begin(), order.end(), [&](std::int64_t a, std::int64_t b) {
        return cells[static_cast<std::size_t>(a)] < cells[static_cast<std::size_t>(b)];
    });
    std::vector<std::int64_t> starts;
    for (std::int64_t q = 0; q < n; ++q)
        if (q == 0 || cells[static_cast<std::size_t>(order[static_cast<std::size_t>(q)])] !=
                          cells[static_cast<std::size_t>(order[static_cast<std::size_t>(q - 1)])])
            starts.push_back(q);
    starts.push_back(n);
    const auto plan = cluster_plan(e, 0.5 * std::sqrt(3.0) * kPointCube, tolerance);
    const std::int64_t nclusters = static_cast<std::int64_t>(starts.size()) - 1;
    for_blocks(nclusters, 1, [&](std::int64_t cidx, int) {
        const std::int64_t b0 = starts[static_cast<std::size_t>(cidx)], b1 = starts[static_cast<std::size_t>(cidx + 1)];
        std::vector<double> pts;
        for (std::int64_t q = b0; q < b1; ++q) {
            const double* r = points + 3 * order[static_cast<std::size_t>(q)];
            pts.insert(pts.end(), r, r + 3);
        }
        const Cell& cell = cells[static_cast<std::size_t>(order[static_cast<std::size_t>(b0)])];
        const double c[3] = {(static_cast<double>(cell[0]) + 0.5) * kPointCube, (static_cast<double>(cell[1]) + 0.5) * kPointCube,
                             (static_cast<double>(cell[2]) + 0.5) * kPointCube};
        std::vector<double> v(static_cast<std::size_t>(b1 - b0));
        FarStats st;
        evaluate_cluster(e, *plan, b1 - b0, pts.data(), c, v.data(), st);
        for (std::int64_t q = b0; q < b1; ++q)
            out[order[static_cast<std::size_t>(q)]] = v[static_cast<std::size_t>(q - b0)];
    });
}

struct Field {
    orc_property key;
    double** out;
};

void evaluate(const Molecule& M, const PointSet& set, const std::vector<orc_property>& props, const orc_options& opt,
              double** outputs, int* backend_used) {
    if (opt.block_size < 1 || opt.block_size > 65536) throw std::invalid_argument("evaluation block_size must be in [1, 65536]");
    if (!std::isfinite(opt.fukui_delta) || opt.fukui_delta <= 0.0)
        throw std::invalid_argument("fukui_delta must be finite and positive");
    if (!std::isfinite(opt.screening_threshold) || opt.screening_threshold < 0.0)
        throw std::invalid_argument("screening threshold must be finite and nonnegative");
    if (!std::isfinite(opt.potential_tolerance) || opt.potential_tolerance < 0.0)
        throw std::invalid_argument("potential tolerance must be finite and nonnegative");
    if (opt.backend == AO_MO && !M.eligible)
        throw std::invalid_argument("AO_MO backend requested but orbital representation is not eligible");
    const Plan plan = make_plan(props);
    const int backend = choose_backend(opt.backend, M.eligible, plan);
    *backend_used = backend;
    const bool grid = set.grid();
    const std::int64_t n = set.n;
    // Requested fields (duplicates evaluated once; they share outputs).
    std::vector<Field> fields;
    for (std::size_t i = 0; i < props.size(); ++i)
        fields.push_back({props[i], outputs + 0});
    {
        // channel offsets into the flat outputs array
        std::size_t off = 0;
        for (std::size_t i = 0; i < props.size(); ++i) {
            fields[i].out = outputs + off;
            off += static_cast<std::size_t>(channels(props[i].kind));
        }
    }
    // The orbital-energy columns, checked before any pass (src/evaluator.cpp:579-596).
    OrbitalEnergySet oes;
    if (plan.orbital_energies) oes = orbital_energy_set(M, plan, opt.fukui_delta);
    // Orbital checks (src/evaluator.cpp:618-632).
    for (const auto& f : fields)
        if (is_orbital_property(f.key.kind)) {
            if (M.nmo(f.key.channel) == 0) throw std::runtime_error("FCHK has no MO coefficients");
            if (f.key.orbital_index > M.nmo(f.key.channel)) throw std::out_of_range("MO index outside available coefficient range");
        }
    const int bs = opt.block_size;
    const double thr = opt.screening_threshold;
    // Point pass (src/evaluator.cpp:634-657).
    std::vector<int> needed;
    for (int c = 0; c < 4; ++c)
        if (uses_point_path(plan.c[static_cast<std::size_t>(c)], backend, grid)) needed.push_back(c);
    if (!needed.empty()) {
        const int ao_order = grid ? plan.max_order : std::max(1, plan.max_order);
        for_blocks(n, bs, [&](std::int64_t start, int size) {
            Scratch s;
            std::vector<double> pts;
            set.block(start, size, pts);
            const BlockDD block = backend == AO_MO ? density_mo(M, pts.data(), size, needed, std::max(0, ao_order), thr, s)
                                                   : density_p(M, pts.data(), size, needed, ao_order, thr, s);
            for (const auto& f : fields)
                if (from_point_path(f.key, backend, grid)) store_block(f.key.kind, f.key.component, block, size, start, f.out);
        });
    }
    // Electrostatic potential pass (src/evaluator.cpp:664-675): target
    // clusters under the far-field plan; with tolerance 0 every pair is exact
    // and the values equal the pointwise sum bit for bit (an empty local
    // expansion adds zero), which is evaluated directly.
    for (const auto& f : fields)
        if (f.key.kind == ElectrostaticPotential) {
            const auto esp = make_esp(M);
            if (opt.potential_tolerance == 0.0) {
                for_blocks(n, 64, [&](std::int64_t start, int size) {
                    double r[3];
                    for (int j = 0; j < size; ++j) {
                        set.point(start + j, r);
                        f.out[0][start + j] = esp_nuclear(*esp, r) + esp_electronic(*esp, r);
                    }
                });
            } else if (grid) {
                esp_grid_clusters(*esp, set, opt.potential_tolerance, f.out[0]);
            } else {
                esp_point_clusters(*esp, n, set.pts, opt.potential_tolerance, f.out[0]);
            }
        }
    // Orbital-energy pass (src/evaluator.cpp:676-682): AO values without
    // per-block screening (the shell cutoff only), the column sums, then the
    // fields of store_orbital_energy_block.
    if (plan.orbital_energies) {
        for_blocks(n, bs, [&](std::int64_t start, int size) {
            Scratch sc;
            std::vector<double> pts;
            set.block(start, size, pts);
            Panels phi;
            ao_panels(*M.basis, size, pts.data(), 0, phi, sc);
            std::vector<double> od, oe, vd, ve, unused, ae, fp, fm;
            column_sums(oes.occupied, phi, M.n, size, od, &oe);
            if (oes.virtuals) {
                column_sums(oes.virt, phi, M.n, size, vd, &ve);
                column_sums(oes.attachment, phi, M.n, size, unused, &ae);
            }
            if (oes.fukui) {
                column_sums(oes.plus, phi, M.n, size, fp, nullptr);
                column_sums(oes.minus, phi, M.n, size, fm, nullptr);
            }
            for (const auto& f : fields) {
                if (!is_orbital_energy(f.key.kind)) continue;
                for (int t = 0; t < size; ++t) {
                    const std::size_t q = static_cast<std::size_t>(t);
                    const double rho = od[q];
                    const double alie = rho > 0.0 ? oe[q] / rho : 0.0;
                    double lea = 0.0, att = 0.0, plus = 0.0, minus = 0.0;
                    if (oes.virtuals) {
                        lea = vd[q] > 0.0 ? -ve[q] / vd[q] : 0.0;
                        att = rho > 0.0 ? ae[q] / rho : 0.0;
                    }
                    if (oes.fukui) {
                        plus = fp[q];
                        minus = fm[q];
                    }
                    double v;
                    switch (f.key.kind) {
                    case AverageLocalIonizationEnergy: v = alie; break;
                    case LocalElectronAffinity: v = lea; break;
                    case LocalElectronAttachmentEnergy: v = att; break;
                    case LocalElectronegativity: v = 0.5 * (alie + lea); break;
                    case LocalHardness: v = 0.5 * (alie - lea); break;
                    case OrbitalWeightedFukuiPlus: v = plus; break;
                    case OrbitalWeightedFukuiMinus: v = minus; break;
                    case OrbitalWeightedFukuiZero: v = 0.5 * (plus + minus); break;
                    default: v = plus - minus; break;
                    }
                    f.out[0][start + t] = v;
                }
            }
        });
    }
    // Orbital pass (src/evaluator.cpp:688-730): derivatives on grids, all at points.
    std::vector<Field> orbital_fields;
    int dorder = 0;
    for (const auto& f : fields)
        if (is_orbital_property(f.key.kind) && !on_grid_path(f.key, backend, grid)) {
            orbital_fields.push_back(f);
            if (f.key.kind == OrbitalGradient) dorder = std::max(dorder, 1);
            if (f.key.kind == OrbitalHessian) dorder = 2;
        }
    if (!orbital_fields.empty()) {
        for_blocks(n, bs, [&](std::int64_t start, int size) {
            Scratch s;
            std::vector<double> pts;
            set.block(start, size, pts);
            Panels phi, pan;
            ao_panels(*M.basis, size, pts.data(), 0, phi, s);
            if (dorder > 0) ao_panels(*M.basis, size, pts.data(), dorder, pan, s);
            for (const auto& f : orbital_fields) {
                const auto& C = M.C(f.key.channel);
                const int nmo = M.nmo(f.key.channel), mo = f.key.orbital_index - 1;
                const auto store = [&](const Panels& X, int panel, int comp) {
                    for (int t = 0; t < size; ++t) {
                        double v = 0.0;
                        for (int a = 0; a < M.n; ++a)
                            v += X.at(a, panel * size + t) * C[static_cast<std::size_t>(a * nmo + mo)];
                        f.out[comp][start + t] = v;
                    }
                };
                if (f.key.kind == Orbital) store(phi, 0, 0);
                else if (f.key.kind == OrbitalGradient) for (int a = 0; a < 3; ++a) store(pan, 1 + a, a);
                else for (int h = 0; h < 6; ++h) store(pan, 4 + h, h);
            }
        });
    }
    // Grid path (src/evaluator.cpp:754-777): AO-P densities, one pass each.
    for (const auto& f : fields) {
        if (!(on_grid_path(f.key, backend, grid) && f.key.kind == Density)) continue;
        const int comp = f.key.component;
        // densities_blocked with one component (src/density.cpp:365-444).
        std::vector<double> Pc;
        const std::vector<double>* Pm = &M.P;
        if (comp == Difference) Pm = &M.S;
        else if (comp == Alpha || comp == Beta) {
            Pc.resize(M.P.size());
            for (std::size_t q = 0; q < Pc.size(); ++q)
                Pc[q] = 0.5 * (comp == Alpha ? (M.P[q] + M.S[q]) : (M.P[q] - M.S[q]));
            Pm = &Pc;
        }
        for_blocks(n, bs, [&](std::int64_t start, int size) {
            Scratch s;
            std::vector<double> pts;
            set.block(start, size, pts);
            Panels phi;
            ao_panels(*M.basis, size, pts.data(), 0, phi, s);
            const auto kept = significant(phi, thr);
            std::vector<double> y;
            p_product(*Pm, M.n, kept, phi, size, y);
            for (int t = 0; t < size; ++t) {
                double r = 0.0;
                for (std::size_t i = 0; i < kept.size(); ++i)
                    r += phi.at(kept[i], t) * y[i * static_cast<std::size_t>(size) + static_cast<std::size_t>(t)];
                f.out[0][start + t] = r;
            }
        });
    }
    // Grid path: orbital amplitudes (src/orbitals.cpp:66-102), blocks of 128, no screening.
    for (const auto& f : fields) {
        if (!(on_grid_path(f.key, backend, grid) && f.key.kind == Orbital)) continue;
        const auto& C = M.C(f.key.channel);
        const int nmo = M.nmo(f.key.channel), mo = f.key.orbital_index - 1;
        for_blocks(n, 128, [&](std::int64_t start, int size) {
            Scratch s;
            std::vector<double> pts;
            set.block(start, size, pts);
            Panels phi;
            ao_panels(*M.basis, size, pts.data(), 0, phi, s);
            for (int t = 0; t < size; ++t) {
                double v = 0.0;
                for (int a = 0; a < M.n; ++a)
                    v += C[static_cast<std::size_t>(a * nmo + mo)] * phi.at(a, t);
                f.out[0][start + t] = v;
            }
        });
    }
}

// ---------------------------------------------------------------------------
// Molecular surfaces: src/surface.cpp (isosurface extraction :150-331, density
// surface :341-369, statistics :377-529).

struct OMesh {
    std::vector<std::array<double, 3>> v;
    std::vector<std::array<std::int64_t, 3>> t;
};

using ScalarFn = std::function<void(const std::vector<std::array<double, 3>>&, std::vector<double>&,
                                    std::vector<std::array<double, 3>>&)>;

struct GridSpec {
    std::array<double, 3> origin;
    std::array<double, 9> steps; // column a = steps[3a .. 3a + 2]
    std::array<int, 3> counts;
    std::array<double, 3> at(std::int64_t q) const {
        PointSet ps;
        ps.origin = origin.data();
        ps.steps = steps.data();
        ps.counts = counts.data();
        std::array<double, 3> r;
        ps.point(q, r.data());
        return r;
    }
    std::int64_t index(int i, int j, int k) const {
        return (static_cast<std::int64_t>(i) * counts[1] + j) * counts[2] + k;
    }
};

bool reaches_boundary(const GridSpec& g, const std::vector<double>& f, double iso) {
    for (int i = 0; i < g.counts[0]; ++i)
        for (int j = 0; j < g.counts[1]; ++j)
            for (int k = 0; k < g.counts[2]; ++k) {
                const bool edge = i == 0 || j == 0 || k == 0 || i == g.counts[0] - 1 || j == g.counts[1] - 1 ||
                                  k == g.counts[2] - 1;
                if (edge && f[static_cast<std::size_t>(g.index(i, j, k))] >= iso) return true;
            }
    return false;
}

std::array<double, 3> sub3(const std::array<double, 3>& a, const std::array<double, 3>& b) {
    return {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
}
std::array<double, 3> cross3(const std::array<double, 3>& a, const std::array<double, 3>& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot3(const std::array<double, 3>& a, const std::array<double, 3>& b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

OMesh isosurface(const GridSpec& g, const std::vector<double>& f, double iso, const ScalarFn& fn) {
    if (g.counts[0] < 2 || g.counts[1] < 2 || g.counts[2] < 2)
        throw std::invalid_argument("isosurface grid needs at least two points along each axis");
    if (f.size() != static_cast<std::size_t>(g.counts[0]) * g.counts[1] * g.counts[2])
        throw std::invalid_argument("isosurface samples do not match the grid");
    if (!std::isfinite(iso)) throw std::invalid_argument("isovalue must be finite");
    for (double x : f)
        if (!std::isfinite(x)) throw std::invalid_argument("isosurface samples must be finite");
    if (reaches_boundary(g, f, iso))
        throw std::invalid_argument("the isosurface reaches the grid boundary, so it would not be closed");
    // Cells in Field order, six tetrahedra each (corners 0, e_a, e_a + e_b, 7
    // for the permutations (a, b, c)); vertices numbered at first use of their
    // edge, the edge named by its lower end and its direction.
    static const int tets[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
    std::map<std::pair<std::int64_t, int>, std::int64_t> edge_vertex;
    std::vector<std::pair<std::int64_t, std::int64_t>> ends; // (inside, outside) per vertex
    std::vector<std::vector<std::int64_t>> polys;
    for (int i = 0; i + 1 < g.counts[0]; ++i)
        for (int j = 0; j + 1 < g.counts[1]; ++j)
            for (int k = 0; k + 1 < g.counts[2]; ++k) {
                std::int64_t idx[8];
                bool in[8];
                int count = 0;
                for (int m = 0; m < 8; ++m) {
                    idx[m] = g.index(i + (m & 1), j + (m >> 1 & 1), k + (m >> 2 & 1));
                    in[m] = f[static_cast<std::size_t>(idx[m])] >= iso;
                    count += in[m] ? 1 : 0;
                }
                if (count == 0 || count == 8) continue;
                for (const auto& tet : tets) {
                    std::vector<int> ins, outs;
                    for (int c = 0; c < 4; ++c) (in[tet[c]] ? ins : outs).push_back(tet[c]);
                    if (ins.empty() || outs.empty()) continue;
                    std::vector<std::pair<int, int>> corners;
                    if (ins.size() == 1)
                        corners = {{ins[0], outs[0]}, {ins[0], outs[1]}, {ins[0], outs[2]}};
                    else if (outs.size() == 1)
                        corners = {{ins[0], outs[0]}, {ins[1], outs[0]}, {ins[2], outs[0]}};
                    else
                        corners = {{ins[0], outs[0]}, {ins[0], outs[1]}, {ins[1], outs[1]}, {ins[1], outs[0]}};
                    std::vector<std::array<double, 3>> mid;
                    for (const auto& [a, b] : corners) {
                        const auto pa = g.at(idx[a]), pb = g.at(idx[b]);
                        mid.push_back({0.5 * (pa[0] + pb[0]), 0.5 * (pa[1] + pb[1]), 0.5 * (pa[2] + pb[2])});
                    }
                    const auto normal = corners.size() == 3 ? cross3(sub3(mid[1], mid[0]), sub3(mid[2], mid[0]))
                                                            : cross3(sub3(mid[2], mid[0]), sub3(mid[3], mid[1]));
                    std::array<double, 3> dir{0, 0, 0};
                    for (int c : outs) {
                        const auto p = g.at(idx[c]);
                        for (int a = 0; a < 3; ++a) dir[a] += p[a] / static_cast<double>(outs.size());
                    }
                    for (int c : ins) {
                        const auto p = g.at(idx[c]);
                        for (int a = 0; a < 3; ++a) dir[a] -= p[a] / static_cast<double>(ins.size());
                    }
                    if (dot3(normal, dir) < 0) std::reverse(corners.begin(), corners.end());
                    std::vector<std::int64_t> poly;
                    for (const auto& [a, b] : corners) {
                        const int lo = std::min(a, b), hi = std::max(a, b);
                        const auto key = std::make_pair(idx[lo], hi ^ lo);
                        auto it = edge_vertex.find(key);
                        if (it == edge_vertex.end()) {
                            it = edge_vertex.emplace(key, static_cast<std::int64_t>(ends.size())).first;
                            ends.push_back({idx[a], idx[b]});
                        }
                        poly.push_back(it->second);
                    }
                    polys.push_back(poly);
                }
            }
    // Vertex placement: t in [0, 1] along (inside, outside), linear start,
    // Newton steps on f - iso kept inside the bracket, else bisection.
    const std::size_t nv = ends.size();
    std::vector<double> t(nv), lo(nv, 0.0), hi(nv, 1.0);
    std::vector<char> done(nv, 0);
    for (std::size_t v = 0; v < nv; ++v) {
        const double a = f[static_cast<std::size_t>(ends[v].first)], b = f[static_cast<std::size_t>(ends[v].second)];
        t[v] = std::min(1.0, std::max(0.0, (a - iso) / (a - b)));
    }
    const auto place = [&](std::size_t v) {
        const auto a = g.at(ends[v].first), b = g.at(ends[v].second);
        return std::array<double, 3>{a[0] + t[v] * (b[0] - a[0]), a[1] + t[v] * (b[1] - a[1]), a[2] + t[v] * (b[2] - a[2])};
    };
    for (int round = 0;; ++round) {
        std::vector<std::size_t> act;
        for (std::size_t v = 0; v < nv; ++v)
            if (!done[v]) act.push_back(v);
        if (act.empty()) break;
        if (round == 200) throw std::runtime_error("isosurface vertices did not converge");
        std::vector<std::array<double, 3>> pts;
        for (std::size_t v : act) pts.push_back(place(v));
        std::vector<double> val(act.size(), 0.0);
        std::vector<std::array<double, 3>> grad(act.size(), {0.0, 0.0, 0.0});
        fn(pts, val, grad);
        for (std::size_t p = 0; p < act.size(); ++p) {
            const std::size_t v = act[p];
            const double r = val[p] - iso;
            if (!std::isfinite(r) || !std::isfinite(grad[p][0]) || !std::isfinite(grad[p][1]) ||
                !std::isfinite(grad[p][2]))
                throw std::runtime_error("isosurface function is not finite");
            if (r == 0.0) {
                done[v] = 1;
                continue;
            }
            if (r > 0.0) lo[v] = t[v];
            else hi[v] = t[v];
            const auto d = sub3(g.at(ends[v].second), g.at(ends[v].first));
            const double len = std::sqrt(dot3(d, d));
            double next = t[v] - r / dot3(grad[p], d);
            if (!(next > lo[v] && next < hi[v])) next = 0.5 * (lo[v] + hi[v]);
            const double moved = std::abs(next - t[v]) * len;
            t[v] = next;
            if (!(moved > 1e-10 && (hi[v] - lo[v]) * len > 1e-10)) done[v] = 1;
        }
    }
    OMesh m;
    for (std::size_t v = 0; v < nv; ++v) m.v.push_back(place(v));
    for (const auto& q : polys) {
        if (q.size() == 3) {
            m.t.push_back({q[0], q[1], q[2]});
            continue;
        }
        const auto d02 = sub3(m.v[static_cast<std::size_t>(q[0])], m.v[static_cast<std::size_t>(q[2])]);
        const auto d13 = sub3(m.v[static_cast<std::size_t>(q[1])], m.v[static_cast<std::size_t>(q[3])]);
        if (dot3(d02, d02) <= dot3(d13, d13)) {
            m.t.push_back({q[0], q[1], q[2]});
            m.t.push_back({q[0], q[2], q[3]});
        } else {
            m.t.push_back({q[0], q[1], q[3]});
            m.t.push_back({q[1], q[2], q[3]});
        }
    }
    return m;
}

// Grid::around (src/grid.cpp:15-38).
GridSpec grid_around(const Molecule& M, double margin, double spacing) {
    std::array<double, 3> lo{M.atoms[0].pos[0], M.atoms[0].pos[1], M.atoms[0].pos[2]}, hi = lo;
    for (const auto& a : M.atoms)
        for (int k = 0; k < 3; ++k) {
            lo[static_cast<std::size_t>(k)] = std::min(lo[static_cast<std::size_t>(k)], a.pos[k]);
            hi[static_cast<std::size_t>(k)] = std::max(hi[static_cast<std::size_t>(k)], a.pos[k]);
        }
    GridSpec g{};
    for (int k = 0; k < 3; ++k) {
        const auto u = static_cast<std::size_t>(k);
        g.origin[u] = lo[u] - margin;
        g.counts[u] = std::max(2, static_cast<int>(std::ceil((hi[u] - lo[u] + 2 * margin) / spacing) + 1));
        g.steps[4 * u] = spacing;
    }
    return g;
}

// The density isosurface: AO-P grid densities (block 128, screening 1e-12)
// and the density with its gradient at the vertices (AO-P, order 1).
OMesh density_isosurface(const Molecule& M, double iso, double spacing) {
    if (!std::isfinite(iso) || iso < 1e-8) throw std::invalid_argument("density isovalue must be finite and at least 1e-8 e/bohr^3");
    if (!std::isfinite(spacing) || spacing <= 0.0) throw std::invalid_argument("surface grid spacing must be positive and finite");
    const orc_options opt{128, AO_P, 1e-12, 0.1, 0.0};
    const ScalarFn fn = [&](const std::vector<std::array<double, 3>>& pts, std::vector<double>& val,
                            std::vector<std::array<double, 3>>& grad) {
        std::vector<double> flat;
        for (const auto& p : pts) flat.insert(flat.end(), p.begin(), p.end());
        PointSet set;
        set.pts = flat.data();
        set.n = static_cast<std::int64_t>(pts.size());
        std::vector<double> gx(pts.size()), gy(pts.size()), gz(pts.size());
        double* outs[4] = {val.data(), gx.data(), gy.data(), gz.data()};
        int used = 0;
        evaluate(M, set, {{Density, Total, 0, 0}, {Gradient, Total, 0, 0}}, opt, outs, &used);
        for (std::size_t p = 0; p < pts.size(); ++p) grad[p] = {gx[p], gy[p], gz[p]};
    };
    for (double margin = 4.0; margin <= 16.0; margin += 2.0) {
        const GridSpec g = grid_around(M, margin, spacing);
        std::vector<double> rho(static_cast<std::size_t>(g.counts[0]) * g.counts[1] * g.counts[2]);
        PointSet set;
        set.origin = g.origin.data();
        set.steps = g.steps.data();
        set.counts = g.counts.data();
        set.n = static_cast<std::int64_t>(rho.size());
        double* out = rho.data();
        int used = 0;
        evaluate(M, set, {{Density, Total, 0, 0}}, opt, &out, &used);
        if (reaches_boundary(g, rho, iso)) continue;
        return isosurface(g, rho, iso, fn);
    }
    throw std::runtime_error("the density isosurface extends more than 16 bohr beyond the atoms; use a larger isovalue");
}

// Politzer-Murray statistics of vertex values on a mesh, from the linear
// interpolation over each triangle.
struct OStats {
    double area = 0, pos_area = 0, neg_area = 0, nonpolar = 0, minimum = 0, maximum = 0;
    std::int64_t min_vertex = 0, max_vertex = 0;
    double average = 0, pos_avg = 0, neg_avg = 0, pos_var = 0, neg_var = 0, tot_var = 0, nu = 0, pi = 0, mpi = 0;
    std::vector<std::int64_t> minima, maxima;
};

// Pieces of a triangle where the linear function >= level (upper) or < level.
void pieces(const std::array<std::array<double, 3>, 3>& x, const std::array<double, 3>& w, double level, bool upper,
            const std::function<void(double, double, double, double)>& piece) {
    std::vector<std::array<double, 3>> px;
    std::vector<double> pw;
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        const bool in_i = upper ? w[static_cast<std::size_t>(i)] >= level : w[static_cast<std::size_t>(i)] < level;
        const bool in_j = upper ? w[static_cast<std::size_t>(j)] >= level : w[static_cast<std::size_t>(j)] < level;
        if (in_i) {
            px.push_back(x[static_cast<std::size_t>(i)]);
            pw.push_back(w[static_cast<std::size_t>(i)]);
        }
        if (in_i != in_j) {
            const double s = (level - w[static_cast<std::size_t>(i)]) / (w[static_cast<std::size_t>(j)] - w[static_cast<std::size_t>(i)]);
            const auto& a = x[static_cast<std::size_t>(i)];
            const auto& b = x[static_cast<std::size_t>(j)];
            px.push_back({a[0] + s * (b[0] - a[0]), a[1] + s * (b[1] - a[1]), a[2] + s * (b[2] - a[2])});
            pw.push_back(level);
        }
    }
    for (std::size_t i = 1; i + 1 < px.size(); ++i) {
        const auto c = cross3(sub3(px[i], px[0]), sub3(px[i + 1], px[0]));
        piece(0.5 * std::sqrt(dot3(c, c)), pw[0], pw[i], pw[i + 1]);
    }
}

OStats statistics(const std::vector<std::array<double, 3>>& xv, const std::vector<std::array<std::int64_t, 3>>& tri,
                  const std::vector<double>& w, double polar) {
    for (double x : w)
        if (!std::isfinite(x)) throw std::invalid_argument("surface values must be finite");
    if (!std::isfinite(polar) || polar < 0.0) throw std::invalid_argument("polar threshold must be finite and nonnegative");
    for (const auto& t : tri)
        for (auto v : t)
            if (v < 0 || v >= static_cast<std::int64_t>(xv.size())) throw std::invalid_argument("surface triangle refers to a missing vertex");
    OStats s;
    const auto corners = [&](const std::array<std::int64_t, 3>& t, std::array<std::array<double, 3>, 3>& x,
                             std::array<double, 3>& v) {
        for (int c = 0; c < 3; ++c) {
            x[static_cast<std::size_t>(c)] = xv[static_cast<std::size_t>(t[static_cast<std::size_t>(c)])];
            v[static_cast<std::size_t>(c)] = w[static_cast<std::size_t>(t[static_cast<std::size_t>(c)])];
        }
    };
    const auto avg3 = [](double a, double b, double c) { return (a + b + c) / 3.0; };
    double ip = 0, in = 0, above = 0, below = 0;
    for (const auto& t : tri) {
        std::array<std::array<double, 3>, 3> x;
        std::array<double, 3> v;
        corners(t, x, v);
        pieces(x, v, 0.0, true, [&](double a, double p, double q, double r) { s.pos_area += a; ip += a * avg3(p, q, r); });
        pieces(x, v, 0.0, false, [&](double a, double p, double q, double r) { s.neg_area += a; in += a * avg3(p, q, r); });
        pieces(x, v, polar, true, [&](double a, double, double, double) { above += a; });
        pieces(x, v, -polar, false, [&](double a, double, double, double) { below += a; });
    }
    s.area = s.pos_area + s.neg_area;
    if (!(s.area > 0.0) || !std::isfinite(s.area)) throw std::invalid_argument("surface area must be positive");
    const double nan = std::numeric_limits<double>::quiet_NaN();
    s.average = (ip + in) / s.area;
    s.pos_avg = s.pos_area > 0 ? ip / s.pos_area : nan;
    s.neg_avg = s.neg_area > 0 ? in / s.neg_area : nan;
    s.mpi = (ip - in) / s.area;
    s.nonpolar = std::max(0.0, s.area - above - below);
    double sp = 0, sn = 0, dev = 0;
    const auto sq = [](double a, double p, double q, double r, double m) {
        const double u = p - m, v = q - m, x = r - m;
        return a / 6.0 * (u * u + v * v + x * x + u * v + v * x + x * u);
    };
    for (const auto& t : tri) {
        std::array<std::array<double, 3>, 3> x;
        std::array<double, 3> v;
        corners(t, x, v);
        if (s.pos_area > 0) pieces(x, v, 0.0, true, [&](double a, double p, double q, double r) { sp += sq(a, p, q, r, s.pos_avg); });
        if (s.neg_area > 0) pieces(x, v, 0.0, false, [&](double a, double p, double q, double r) { sn += sq(a, p, q, r, s.neg_avg); });
        pieces(x, v, s.average, true, [&](double a, double p, double q, double r) { dev += a * (avg3(p, q, r) - s.average); });
        pieces(x, v, s.average, false, [&](double a, double p, double q, double r) { dev += a * (s.average - avg3(p, q, r)); });
    }
    s.pos_var = s.pos_area > 0 ? sp / s.pos_area : 0.0;
    s.neg_var = s.neg_area > 0 ? sn / s.neg_area : 0.0;
    s.tot_var = s.pos_var + s.neg_var;
    s.nu = s.tot_var > 0 ? s.pos_var * s.neg_var / (s.tot_var * s.tot_var) : nan;
    s.pi = dev / s.area;
    s.min_vertex = std::min_element(w.begin(), w.end()) - w.begin();
    s.max_vertex = std::max_element(w.begin(), w.end()) - w.begin();
    s.minimum = w[static_cast<std::size_t>(s.min_vertex)];
    s.maximum = w[static_cast<std::size_t>(s.max_vertex)];
    // Extrema: strictly below (above) every vertex within two edges.
    std::vector<std::set<std::int64_t>> nb(xv.size());
    for (const auto& t : tri)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (a != b) nb[static_cast<std::size_t>(t[static_cast<std::size_t>(a)])].insert(t[static_cast<std::size_t>(b)]);
    for (std::size_t i = 0; i < xv.size(); ++i) {
        if (nb[i].empty()) continue;
        bool is_min = true, is_max = true;
        for (auto j : nb[i]) {
            std::vector<std::int64_t> ring{j};
            ring.insert(ring.end(), nb[static_cast<std::size_t>(j)].begin(), nb[static_cast<std::size_t>(j)].end());
            for (auto k : ring) {
                if (k == static_cast<std::int64_t>(i)) continue;
                if (!(w[i] < w[static_cast<std::size_t>(k)])) is_min = false;
                if (!(w[i] > w[static_cast<std::size_t>(k)])) is_max = false;
            }
        }
        if (is_min) s.minima.push_back(static_cast<std::int64_t>(i));
        if (is_max) s.maxima.push_back(static_cast<std::int64_t>(i));
    }
    std::stable_sort(s.minima.begin(), s.minima.end(), [&](std::int64_t a, std::int64_t b) { return w[static_cast<std::size_t>(a)] < w[static_cast<std::size_t>(b)]; });
    std::stable_sort(s.maxima.begin(), s.maxima.end(), [&](std::int64_t a, std::int64_t b) { return w[static_cast<std::size_t>(a)] > w[static_cast<std::size_t>(b)]; });
    return s;
}

void set_error(char* buf, int len, const char* msg) {
    if (buf && len > 0) {
        std::strncpy(buf, msg, static_cast<std::size_t>(len - 1));
        buf[len - 1] = '\0';
    }
}

} // namespace orc

using namespace orc;

extern "C" {

void* orc_basis_create(int nshell, const int* shell_types, const int* nprim, const int* shell_atom,
                       const double* shell_xyz, const double* exponents, const double* coefficients,
                       const double* p_coefficients, char* error, int error_len) {
    try {
        return make_basis(nshell, shell_types, nprim, shell_atom, shell_xyz, exponents, coefficients, p_coefficients)
            .release();
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        return nullptr;
    }
}

void orc_basis_destroy(void* b) { delete static_cast<Basis*>(b); }
int orc_basis_nao(const void* b) { return static_cast<const Basis*>(b)->nao; }
int orc_basis_nblocks(const void* b) { return static_cast<int>(static_cast<const Basis*>(b)->blocks.size()); }

void orc_basis_blocks(const void* b, int* l, int* offset, int* count, double* cutoff_r2) {
    const auto* B = static_cast<const Basis*>(b);
    for (std::size_t i = 0; i < B->blocks.size(); ++i) {
        l[i] = B->blocks[i].l;
        offset[i] = B->blocks[i].offset;
        count[i] = B->blocks[i].count;
        cutoff_r2[i] = B->blocks[i].cutoff_r2;
    }
}

int orc_ao_eval(const void* b, int n, const double* points, int order, double* out) {
    if (order < 0 || order > 2) return 1;
    const auto* B = static_cast<const Basis*>(b);
    Panels p;
    Scratch s;
    ao_panels(*B, n, points, order, p, s);
    // column-major output: (row, col) at row + nao * col
    for (int r = 0; r < p.rows; ++r)
        for (int c = 0; c < p.cols; ++c)
            out[static_cast<std::size_t>(r) + static_cast<std::size_t>(p.rows) * static_cast<std::size_t>(c)] = p.at(r, c);
    return 0;
}

void orc_overlap(const void* b, double* out) {
    const auto S = overlap(*static_cast<const Basis*>(b));
    std::copy(S.begin(), S.end(), out);
}

int orc_pure_transform(int l, double* out) {
    if (l < 2) return 1;
    const auto t = pure_transform(l);
    std::copy(t.begin(), t.end(), out);
    return 0;
}

void* orc_molecule_create(void* basis, int natoms, const int* atomic_numbers, const double* charges,
                          const double* xyz, const double* p_total, const double* p_spin, int nmo_alpha,
                          const double* c_alpha, int nmo_beta, const double* c_beta, int nalpha, int nbeta,
                          char* error, int error_len) {
    try {
        auto m = std::make_unique<Molecule>();
        m->basis = static_cast<const Basis*>(basis);
        m->n = m->basis->nao;
        const std::size_t n = static_cast<std::size_t>(m->n);
        for (int a = 0; a < natoms; ++a)
            m->atoms.push_back({atomic_numbers[a], charges[a], {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]}});
        m->P.assign(p_total, p_total + n * n);
        if (p_spin) m->S.assign(p_spin, p_spin + n * n);
        else m->S.assign(n * n, 0.0);
        // MO-major (AO fastest) -> AO rows x MO columns.
        const auto load = [&](const double* src, int nmo, std::vector<double>& dst) {
            dst.assign(n * static_cast<std::size_t>(nmo), 0.0);
            for (int mo = 0; mo < nmo; ++mo)
                for (std::size_t ao = 0; ao < n; ++ao)
                    dst[ao * static_cast<std::size_t>(nmo) + static_cast<std::size_t>(mo)] = src[static_cast<std::size_t>(mo) * n + ao];
        };
        if (c_alpha && nmo_alpha > 0) {
            load(c_alpha, nmo_alpha, m->Ca);
            m->nmo_a = nmo_alpha;
        }
        if (c_beta && nmo_beta > 0) {
            load(c_beta, nmo_beta, m->Cb);
            m->nmo_b = nmo_beta;
        }
        m->unrestricted = m->nmo_b != 0;
        m->nalpha = nalpha;
        m->nbeta = nbeta;
        if (nalpha != nbeta && !m->unrestricted) throw std::runtime_error("open-shell FCHK lacks Beta MO coefficients");
        m->eligible = check_eligible(*m);
        return m.release();
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        return nullptr;
    }
}

void orc_molecule_destroy(void* m) { delete static_cast<Molecule*>(m); }
void orc_molecule_set_energies(void* m, int n_alpha, const double* e_alpha, int n_beta, const double* e_beta) {
    auto* M = static_cast<Molecule*>(m);
    M->Ea.assign(e_alpha, e_alpha + (e_alpha ? n_alpha : 0));
    M->Eb.assign(e_beta, e_beta + (e_beta ? n_beta : 0));
}
int orc_molecule_eligible(const void* m) { return static_cast<const Molecule*>(m)->eligible ? 1 : 0; }

static PointSet make_set(const double* origin, const double* steps, const int* counts, std::int64_t npoints,
                         const double* points) {
    PointSet set;
    if (points) {
        set.pts = points;
        set.n = npoints;
    } else {
        set.origin = origin;
        set.steps = steps;
        set.counts = counts;
        set.n = static_cast<std::int64_t>(counts[0]) * counts[1] * counts[2];
    }
    return set;
}

int orc_evaluate(const void* molecule, const double* origin, const double* steps, const int* counts,
                 std::int64_t npoints, const double* points, const orc_property* props, int nprops,
                 const orc_options* options, double** outputs, int* backend_used, char* error, int error_len) {
    try {
        const auto* M = static_cast<const Molecule*>(molecule);
        const PointSet set = make_set(origin, steps, counts, npoints, points);
        std::vector<orc_property> p(props, props + nprops);
        evaluate(*M, set, p, *options, outputs, backend_used);
        return 0;
    } catch (const std::invalid_argument& e) {
        set_error(error, error_len, e.what());
        return 1;
    } catch (const std::out_of_range& e) {
        set_error(error, error_len, e.what());
        return 2;
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        return 3;
    }
}

int orc_masks(const void* molecule, const double* origin, const double* steps, const int* counts,
              std::int64_t npoints, const double* points, int block_size, int ao_order, double threshold,
              std::uint8_t* kept) {
    const auto* M = static_cast<const Molecule*>(molecule);
    const PointSet set = make_set(origin, steps, counts, npoints, points);
    const int nao = M->n;
    for_blocks(set.n, block_size, [&](std::int64_t start, int size) {
        Scratch s;
        std::vector<double> pts;
        set.block(start, size, pts);
        Panels ao;
        ao_panels(*M->basis, size, pts.data(), ao_order, ao, s);
        const auto rows = significant(ao, threshold);
        std::uint8_t* dst = kept + static_cast<std::size_t>(start / block_size) * static_cast<std::size_t>(nao);
        std::fill(dst, dst + nao, 0);
        for (int r : rows)
            dst[r] = 1;
    });
    return 0;
}

int orc_boys(double T, int m, double* out) {
    if (m < 0 || m > kBoysMax || !std::isfinite(T) || T < 0.0) return 1;
    boys(T, m, out);
    return 0;
}

int orc_esp(const void* molecule, std::int64_t n, const double* points, double* out) {
    try {
        const auto esp = make_esp(*static_cast<const Molecule*>(molecule));
        for_blocks(n, 64, [&](std::int64_t start, int size) {
            for (int j = 0; j < size; ++j)
                out[start + j] = esp_nuclear(*esp, points + 3 * (start + j)) + esp_electronic(*esp, points + 3 * (start + j));
        });
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

std::int64_t orc_esp_pair_count(const void* molecule) {
    return static_cast<std::int64_t>(make_esp(*static_cast<const Molecule*>(molecule))->pairs.size());
}

void* orc_esp_create(const void* molecule, char* error, int error_len) {
    try {
        return make_esp(*static_cast<const Molecule*>(molecule)).release();
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        return nullptr;
    }
}

void orc_esp_destroy(void* esp) { delete static_cast<Esp*>(esp); }

void* orc_far_plan(const void* esp, double radius, double tolerance, char* error, int error_len) {
    try {
        return far_field_plan(*static_cast<const Esp*>(esp), radius, tolerance).release();
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        return nullptr;
    }
}

void orc_far_plan_destroy(void* plan) { delete static_cast<FarPlan*>(plan); }

void orc_far_plan_info(const void* plan, std::int64_t* sizes, double* values) {
    const FarPlan& p = *static_cast<const FarPlan*>(plan);
    sizes[0] = static_cast<std::int64_t>(p.npairs);
    sizes[1] = std::count_if(p.level_d2.begin(), p.level_d2.end(), [](float d2) { return std::isfinite(d2); });
    sizes[2] = static_cast<std::int64_t>(p.clusters.size());
    sizes[3] = static_cast<std::int64_t>(p.moment_count);
    values[0] = p.pair_tolerance;
    values[1] = p.radius;
}

int orc_far_plan_cluster(const void* plan, std::int64_t c, double* centre, double* radius, double* min_distance,
                         std::int64_t* npairs, double* moments) {
    const FarPlan& p = *static_cast<const FarPlan*>(plan);
    if (c < 0 || c >= static_cast<std::int64_t>(p.clusters.size())) return 1;
    const SourceCluster& cl = p.clusters[static_cast<std::size_t>(c)];
    for (int a = 0; a < 3; ++a)
        centre[a] = cl.c[a];
    *radius = cl.radius;
    *min_distance = cl.min_distance;
    *npairs = static_cast<std::int64_t>(cl.pairs.size());
    if (moments)
        std::copy(cl.moments.end() - static_cast<std::ptrdiff_t>(p.moment_count), cl.moments.end(), moments);
    return 0;
}

int orc_far_plan_levels(const void* plan, std::int64_t i, float* d2) {
    const FarPlan& p = *static_cast<const FarPlan*>(plan);
    if (i < 0 || i >= static_cast<std::int64_t>(p.npairs)) return 1;
    std::copy_n(p.level_d2.begin() + i * kLevelCount, kLevelCount, d2);
    return 0;
}

int orc_far_cluster(const void* esp, const void* plan, std::int64_t n, const double* points, const double* centre,
                    double* out, std::int64_t* stats, char* error, int error_len) {
    try {
        FarStats st;
        evaluate_cluster(*static_cast<const Esp*>(esp), *static_cast<const FarPlan*>(plan), n, points, centre, out, st);
        stats[0] = st.far_clusters;
        stats[1] = st.expanded;
        stats[2] = st.near;
        return 0;
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        return 1;
    }
}

void orc_grid_points(const double* origin, const double* steps, const int* counts, std::int64_t start, std::int64_t n,
                     double* out) {
    PointSet set = make_set(origin, steps, counts, 0, nullptr);
    for (std::int64_t q = 0; q < n; ++q)
        set.point(start + q, out + 3 * q);
}

// Field::write_cube / write_orbital_cube (src/cube.cpp:70-111, 253-262):
// the samples must be finite (checked before the file is opened), then the
// header with std::fixed / setprecision(6) (atom count width 5, negative for
// an orbital Cube; origin, steps, charges and positions width 13), and the
// samples with std::scientific / setprecision(5) / setw(13), a line break
// after every sixth sample of the whole sequence and after the last sample of
// every run along the third axis.
int orc_write_field_cube(const char* path, const char* comment, const double* origin, const double* steps,
                         const int* counts, int natoms, const int* z, const double* charges, const double* xyz,
                         const double* values, std::int64_t n, int orbital, std::int64_t orbital_index, char* error,
                         int error_len) {
    if (orbital && (orbital_index <= 0 || orbital_index > 9999)) {
        set_error(error, error_len, "invalid orbital index for Cube");
        return 1;
    }
    for (std::int64_t i = 0; i < n; ++i)
        if (!std::isfinite(values[i])) {
            set_error(error, error_len, "nonfinite Cube sample");
            return 1;
        }
    std::ofstream out(path);
    if (!out) {
        set_error(error, error_len, orbital ? "cannot open orbital cube output" : "cannot open cube output");
        return 3;
    }
    out << "chem-kumo 2.1.0\n" << comment << '\n' << std::fixed << std::setprecision(6);
    out << std::setw(5) << static_cast<long long>(orbital ? -natoms : natoms);
    for (int k = 0; k < 3; ++k) out << std::setw(13) << origin[k];
    out << '\n';
    for (int a = 0; a < 3; ++a) {
        out << std::setw(5) << counts[a];
        for (int k = 0; k < 3; ++k) out << std::setw(13) << steps[3 * a + k];
        out << '\n';
    }
    for (int i = 0; i < natoms; ++i) {
        out << std::setw(5) << z[i] << std::setw(13) << charges[i];
        for (int k = 0; k < 3; ++k) out << std::setw(13) << xyz[3 * i + k];
        out << '\n';
    }
    if (orbital) out << std::setw(5) << 1 << std::setw(5) << orbital_index << '\n';
    out << std::scientific << std::setprecision(5);
    const std::int64_t run = counts[2];
    for (std::int64_t i = 1; i <= n; ++i) {
        out << std::setw(13) << values[i - 1];
        if (i % 6 == 0 || i % run == 0) out << '\n';
    }
    if (!out) {
        set_error(error, error_len, orbital ? "orbital cube write failed" : "cube write failed");
        return 3;
    }
    return 0;
}

void* orc_isosurface(const double* origin, const double* steps, const int* counts, const double* samples,
                     double isovalue, orc_scalar_fn fn, void* user, int* status, char* error, int error_len) {
    *status = 0;
    try {
        GridSpec g{};
        for (int k = 0; k < 3; ++k) {
            g.origin[static_cast<std::size_t>(k)] = origin[k];
            g.counts[static_cast<std::size_t>(k)] = counts[k];
        }
        for (int k = 0; k < 9; ++k) g.steps[static_cast<std::size_t>(k)] = steps[k];
        const std::size_t n = static_cast<std::size_t>(counts[0]) * counts[1] * counts[2];
        const std::vector<double> f(samples, samples + n);
        const ScalarFn call = [&](const std::vector<std::array<double, 3>>& pts, std::vector<double>& val,
                                  std::vector<std::array<double, 3>>& grad) {
            std::vector<double> flat, g3(3 * pts.size());
            for (const auto& p : pts) flat.insert(flat.end(), p.begin(), p.end());
            fn(flat.data(), static_cast<std::int64_t>(pts.size()), val.data(), g3.data(), user);
            for (std::size_t q = 0; q < pts.size(); ++q) grad[q] = {g3[3 * q], g3[3 * q + 1], g3[3 * q + 2]};
        };
        return new OMesh(isosurface(g, f, isovalue, call));
    } catch (const std::invalid_argument& e) {
        set_error(error, error_len, e.what());
        *status = 1;
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        *status = 3;
    }
    return nullptr;
}

void* orc_density_isosurface(const void* molecule, double isovalue, double spacing, int* status, char* error,
                             int error_len) {
    *status = 0;
    try {
        return new OMesh(density_isosurface(*static_cast<const Molecule*>(molecule), isovalue, spacing));
    } catch (const std::invalid_argument& e) {
        set_error(error, error_len, e.what());
        *status = 1;
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        *status = 3;
    }
    return nullptr;
}

void orc_mesh_sizes(const void* mesh, std::int64_t* nv, std::int64_t* nt) {
    const auto* m = static_cast<const OMesh*>(mesh);
    *nv = static_cast<std::int64_t>(m->v.size());
    *nt = static_cast<std::int64_t>(m->t.size());
}

void orc_mesh_data(const void* mesh, double* vertices, std::int64_t* triangles) {
    const auto* m = static_cast<const OMesh*>(mesh);
    for (std::size_t v = 0; v < m->v.size(); ++v)
        for (int k = 0; k < 3; ++k) vertices[3 * v + static_cast<std::size_t>(k)] = m->v[v][static_cast<std::size_t>(k)];
    for (std::size_t t = 0; t < m->t.size(); ++t)
        for (int k = 0; k < 3; ++k) triangles[3 * t + static_cast<std::size_t>(k)] = m->t[t][static_cast<std::size_t>(k)];
}

void orc_mesh_free(void* mesh) { delete static_cast<OMesh*>(mesh); }

int orc_surface_statistics(const double* vertices, std::int64_t nv, const std::int64_t* triangles, std::int64_t nt,
                           const double* values, double polar, double* stats, std::int64_t* counts,
                           std::int64_t* minima, std::int64_t* maxima, char* error, int error_len) {
    try {
        std::vector<std::array<double, 3>> xv(static_cast<std::size_t>(nv));
        for (std::int64_t v = 0; v < nv; ++v) xv[static_cast<std::size_t>(v)] = {vertices[3 * v], vertices[3 * v + 1], vertices[3 * v + 2]};
        std::vector<std::array<std::int64_t, 3>> tri(static_cast<std::size_t>(nt));
        for (std::int64_t t = 0; t < nt; ++t) tri[static_cast<std::size_t>(t)] = {triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]};
        const OStats s = statistics(xv, tri, std::vector<double>(values, values + nv), polar);
        const double out[17] = {s.area, s.pos_area, s.neg_area, s.nonpolar, s.minimum, s.maximum,
                                static_cast<double>(s.min_vertex), static_cast<double>(s.max_vertex), s.average,
                                s.pos_avg, s.neg_avg, s.pos_var, s.neg_var, s.tot_var, s.nu, s.pi, s.mpi};
        std::copy(out, out + 17, stats);
        counts[0] = static_cast<std::int64_t>(s.minima.size());
        counts[1] = static_cast<std::int64_t>(s.maxima.size());
        if (minima) std::copy(s.minima.begin(), s.minima.end(), minima);
        if (maxima) std::copy(s.maxima.begin(), s.maxima.end(), maxima);
        return 0;
    } catch (const std::invalid_argument& e) {
        set_error(error, error_len, e.what());
        return 1;
    } catch (const std::exception& e) {
        set_error(error, error_len, e.what());
        return 3;
    }
}

int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}
}
