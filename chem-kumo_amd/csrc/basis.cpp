// Host-side basis preparation (see basis.hpp).
#include "basis.hpp"

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <string>

namespace kumo {
namespace {

double factorial(int n) {
    double v = 1.0;
    for (long long i = 2; i <= n; ++i)
        v *= static_cast<double>(i);
    return v;
}

// Schlegel & Frisch, Int. J. Quantum Chem. 54, 83 (1995): coefficient of
// x^lx y^ly z^lz in the real pure Gaussian (l, m). The factor grouping is the
// one of src/basis_evaluator.cpp:49-83, so that the transform agrees with the
// reference to the last bit.
double pure_coefficient(int l, int m, int lx, int ly) {
    const int lz = l - lx - ly, ma = std::abs(m);
    if (ma > l || lz < 0) return 0.0;
    const double a = std::sqrt(factorial(2 * lx) * factorial(2 * ly) * factorial(2 * lz) * factorial(l) /
                               (factorial(2 * l) * factorial(lx) * factorial(ly) * factorial(lz)));
    const double b = std::sqrt(factorial(l - ma) / factorial(l + ma)) / (std::pow(2.0, l) * factorial(l));
    const int half = (l - ma) / 2;
    const int j = (l - ma - lz) / 2;
    if (2 * j != l - ma - lz) return 0.0;
    double g = 0.0;
    for (int i = 0; i <= half; ++i) {
        if (j < 0 || j > i) continue;
        const double c = factorial(l) / (factorial(i) * factorial(l - i)) * factorial(2 * l - 2 * i) *
                         std::pow(-1.0, i) / factorial(l - ma - 2 * i) * factorial(i) /
                         (factorial(j) * factorial(i - j));
        for (int k = 0; k <= j; ++k) {
            const int px = lx - 2 * k;
            if (px < 0 || px > ma) continue;
            const double d = factorial(j) / (factorial(k) * factorial(j - k)) * factorial(ma) /
                             (factorial(px) * factorial(ma + 2 * k - lx));
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

double odd_double_factorial(int n) {
    double v = 1.0;
    for (int i = n; i >= 1; i -= 2)
        v *= static_cast<double>(i);
    return v;
}

// src/basis_math.cpp:17-25.
double primitive_norm(double alpha, const CartPower& p) {
    const int l = p.x + p.y + p.z;
    const double radial = std::pow(2.0 * alpha / std::numbers::pi, 0.75);
    const double angular = std::sqrt(std::pow(4.0 * alpha, static_cast<double>(l)) /
                                     (odd_double_factorial(2 * p.x - 1) * odd_double_factorial(2 * p.y - 1) *
                                      odd_double_factorial(2 * p.z - 1)));
    return radial * angular;
}

// Rigorous cutoff radius (src/basis_evaluator.cpp:98-127): beyond it every AO
// of the block and its first and second derivatives are at most `cutoff`.
double cutoff_radius(const HostBlock& b, double tnorm, double cutoff) {
    const int l = b.l, np = b.nprim();
    std::vector<double> coef(static_cast<std::size_t>(np));
    double start = 1.0;
    for (int p = 0; p < np; ++p) {
        const double a = b.exps[static_cast<std::size_t>(p)];
        double w = 0.0;
        for (int k = 0; k < b.ncart(); ++k)
            w = std::max(w, std::abs(b.weights[static_cast<std::size_t>(k * np + p)]));
        coef[static_cast<std::size_t>(p)] = tnorm * w * (1.0 + l + l * l + 2.0 * a * (2.0 * l + 2.0) + 4.0 * a * a);
        start = std::max(start, std::sqrt((l + 2.0) / (2.0 * a)));
    }
    const auto bound = [&](double r) {
        double s = 0.0;
        for (int p = 0; p < np; ++p)
            s += coef[static_cast<std::size_t>(p)] * std::pow(r, l + 2) * std::exp(-b.exps[static_cast<std::size_t>(p)] * r * r);
        return s;
    };
    if (bound(start) <= cutoff) return start;
    double lo = start, hi = 2.0 * start;
    while (bound(hi) > cutoff)
        hi *= 2.0;
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        (bound(mid) > cutoff ? lo : hi) = mid;
    }
    return hi;
}

// Largest row sum of |T| (1 for Cartesian blocks), the factor the cutoff
// bound carries for pure AOs (src/basis_evaluator.cpp:98-127).
double transform_norm(const HostBlock& b) {
    if (b.pure.empty()) return 1.0;
    const int nc = b.ncart();
    double tnorm = 0.0;
    for (int row = 0; row < b.count; ++row) {
        double sum = 0.0;
        for (int k = 0; k < nc; ++k)
            sum += std::abs(b.pure[static_cast<std::size_t>(row * nc + k)]);
        tnorm = std::max(tnorm, sum);
    }
    return tnorm;
}

} // namespace

std::vector<double> screening_radii2(const PreparedBasis& basis, double threshold) {
    std::vector<double> r2(basis.blocks.size());
    for (std::size_t i = 0; i < basis.blocks.size(); ++i) {
        const HostBlock& b = basis.blocks[i];
        if (!(threshold > 1e-14)) {
            r2[i] = b.cutoff_r2;
            continue;
        }
        const double r = cutoff_radius(b, transform_norm(b), threshold);
        r2[i] = std::min(b.cutoff_r2, r * r);
    }
    return r2;
}

std::vector<CartPower> gaussian_cartesian_order(int l) {
    if (l == 0) return {{0, 0, 0}};
    if (l == 1) return {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (l == 2) return {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    if (l == 3)
        return {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0},
                {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}};
    std::vector<CartPower> out;
    for (int lx = 0; lx <= l; ++lx)
        for (int ly = 0; ly <= l - lx; ++ly)
            out.push_back({lx, ly, l - lx - ly});
    return out;
}

std::vector<double> gaussian_cartesian_to_pure(int l) {
    const auto order = gaussian_cartesian_order(l);
    const int nc = static_cast<int>(order.size());
    std::vector<double> t(static_cast<std::size_t>((2 * l + 1) * nc), 0.0);
    for (int g = 0; g < nc; ++g) {
        const auto& p = order[static_cast<std::size_t>(g)];
        t[static_cast<std::size_t>(g)] = pure_coefficient(l, 0, p.x, p.y);
        for (int m = 1; m <= l; ++m) {
            t[static_cast<std::size_t>((2 * m - 1) * nc + g)] = pure_coefficient(l, m, p.x, p.y);
            t[static_cast<std::size_t>(2 * m * nc + g)] = pure_coefficient(l, -m, p.x, p.y);
        }
    }
    return t;
}

double component_factor(const CartPower& first, const CartPower& p) {
    const double d0 = odd_double_factorial(2 * first.x - 1) * odd_double_factorial(2 * first.y - 1) *
                      odd_double_factorial(2 * first.z - 1);
    const double dk = odd_double_factorial(2 * p.x - 1) * odd_double_factorial(2 * p.y - 1) *
                      odd_double_factorial(2 * p.z - 1);
    return std::sqrt(d0 / dk);
}

PreparedBasis prepare_basis(const ShellRecords& r) {
    PreparedBasis out;
    if (r.nshells <= 0) runtime("fchk basis has no shells");
    std::size_t prim = 0;
    int offset = 0;
    bool has_sp = false;
    for (int s = 0; s < r.nshells; ++s)
        has_sp = has_sp || r.types[s] == -1;
    if (has_sp && !r.pcoef) runtime("missing P(S=P) coefficients for sp shell");
    const auto add = [&](int s, int l, const double* c, bool pure) {
        if (l > kMaxL) runtime("angular momentum above " + std::to_string(kMaxL) + " is not supported by the HIP backend");
        HostBlock b;
        for (int a = 0; a < 3; ++a)
            b.origin[a] = r.xyz[3 * s + a];
        b.l = l;
        b.atom = r.atom[s] - 1;
        b.exps.assign(r.exps + prim, r.exps + prim + r.nprim[s]);
        b.powers = gaussian_cartesian_order(l);
        const int nc = b.ncart(), np = r.nprim[s];
        b.weights.resize(static_cast<std::size_t>(nc * np));
        for (int k = 0; k < nc; ++k)
            for (int p = 0; p < np; ++p)
                b.weights[static_cast<std::size_t>(k * np + p)] =
                    c[prim + static_cast<std::size_t>(p)] *
                    primitive_norm(b.exps[static_cast<std::size_t>(p)], b.powers[static_cast<std::size_t>(k)]);
        if (pure && l >= 2) b.pure = gaussian_cartesian_to_pure(l);
        b.offset = offset;
        b.count = b.pure.empty() ? nc : 2 * l + 1;
        offset += b.count;
        const double radius = cutoff_radius(b, transform_norm(b), 1e-14);
        b.cutoff_r2 = radius * radius;
        out.max_l = std::max(out.max_l, l);
        out.blocks.push_back(std::move(b));
    };
    for (int s = 0; s < r.nshells; ++s) {
        const int t = r.types[s];
        const int np = r.nprim[s];
        if (np <= 0 || np > kMaxPrim) runtime("invalid primitive count in fchk shell");
        if (r.atom[s] <= 0 || r.atom[s] > r.natoms) runtime("invalid 1-based shell center in fchk");
        for (int p = 0; p < np; ++p) {
            const double e = r.exps[prim + static_cast<std::size_t>(p)];
            if (!(e > 0.0) || !std::isfinite(e) || !std::isfinite(r.coef[prim + static_cast<std::size_t>(p)]) ||
                (r.pcoef && !std::isfinite(r.pcoef[prim + static_cast<std::size_t>(p)])))
                runtime("invalid primitive exponent/coefficient in fchk shell");
        }
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(r.xyz[3 * s + a])) runtime("non-finite shell coordinate in fchk");
        if (t == -1) {
            add(s, 0, r.coef, false);
            add(s, 1, r.pcoef, false);
        } else {
            add(s, t < 0 ? -t : t, r.coef, t <= -2);
        }
        prim += static_cast<std::size_t>(np);
    }
    out.nao = offset;
    if (r.nbasis && out.nao != r.nbasis) runtime("basis size reconstructed from shells does not match fchk");

    // Device tables.
    for (int l = 0; l <= kMaxL + 1; ++l)
        out.pure_index.first[l] = 0;
    for (int l = 0; l <= kMaxL; ++l) {
        out.pure_index.first[l] = static_cast<int32_t>(out.pure_terms.size());
        if (l < 2) continue;
        const auto t = gaussian_cartesian_to_pure(l);
        const int nc = (l + 1) * (l + 2) / 2;
        for (int k = 0; k < nc; ++k)
            for (int row = 0; row < 2 * l + 1; ++row) {
                const double c = t[static_cast<std::size_t>(row * nc + k)];
                if (c != 0.0) out.pure_terms.push_back({c, k, row});
            }
    }
    out.pure_index.first[kMaxL + 1] = static_cast<int32_t>(out.pure_terms.size());
    int cart_rows = 0;
    for (const auto& b : out.blocks)
        cart_rows += b.ncart();
    for (const auto& b : out.blocks) {
        DevShell d{};
        for (int a = 0; a < 3; ++a)
            d.o[a] = b.origin[a];
        d.cutoff_r2 = b.cutoff_r2;
        d.l = b.l;
        d.nprim = b.nprim();
        d.prim = static_cast<int32_t>(out.exps.size());
        d.wofs = static_cast<int32_t>(out.weights.size());
        d.ncart = b.ncart();
        d.count = b.count;
        d.ao = b.offset;
        d.pure = b.pure.empty() ? 0 : 1;
        d.atom = b.atom;
        out.exps.insert(out.exps.end(), b.exps.begin(), b.exps.end());
        out.weights.insert(out.weights.end(), b.weights.begin(), b.weights.end());
        // Factored weights: weights[k][p] = g[k] * weights[0][p].
        const int np = b.nprim(), nc = b.ncart();
        for (int p = 0; p < np; ++p) {
            const double w0 = b.weights[static_cast<std::size_t>(p)];
            out.radw.push_back(w0);
            for (int k = 0; k < nc; ++k) {
                const double wk = b.weights[static_cast<std::size_t>(k * np + p)];
                const double gk = component_factor(b.powers[0], b.powers[static_cast<std::size_t>(k)]);
                if (std::abs(gk * w0 - wk) > 1e-13 * std::abs(wk))
                    runtime("contraction weights of a shell do not factor into radial and angular parts");
            }
        }
        out.shells.push_back(d);
    }
    int crow = 0, pure_rows = 0;
    for (std::size_t i = 0; i < out.shells.size(); ++i) {
        DevShell& d = out.shells[i];
        d.crow = crow;
        crow += d.ncart;
        d.prow = d.pure ? cart_rows + pure_rows : -1;
        if (d.pure) pure_rows += d.count;
    }
    out.operand_rows = cart_rows + pure_rows;
    for (int l = kMaxL; l >= 0; --l)
        for (std::size_t i = 0; i < out.shells.size(); ++i)
            if (out.shells[i].l == l) out.lorder.push_back(static_cast<int32_t>(i));
    return out;
}

} // namespace kumo
