// Host planning of the far-field ESP (src/electrostatic_potential.cpp:363-699).
#include "far_field.hpp"

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <numbers>

namespace kumo {
namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

struct Tuv {
    int t, u, v;
};
// (t, u, v) with t + u + v <= L: by total degree, then t and u descending
// (the order of the pairs' Hermite coefficients).
std::vector<Tuv> tuv_list(int L) {
    std::vector<Tuv> out;
    for (int s = 0; s <= L; ++s)
        for (int t = s; t >= 0; --t)
            for (int u = s - t; u >= 0; --u)
                out.push_back({t, u, s - t - u});
    return out;
}

// Weight of a stored Hermite coefficient (which carries 2 pi / p) as a point
// multipole: sqrt(pi / p) / 2 (src/electrostatic_potential.cpp:31-33).
double multipole_factor(double p) { return 0.5 * std::sqrt(std::numbers::pi / p); }

// W_o: the sum of |e_k| over the pair's Hermite terms of total order o
// (src/electrostatic_potential.cpp:484-495).
std::vector<double> order_weights(const EspPairDev& pr, const double* c) {
    std::vector<double> w(static_cast<std::size_t>(pr.L + 1), 0.0);
    const double f = multipole_factor(pr.p);
    int k = 0;
    for (int o = 0; o <= pr.L; ++o)
        for (int j = 0; j < (o + 1) * (o + 2) / 2; ++j, ++k)
            w[static_cast<std::size_t>(o)] += std::fabs(c[k]) * f;
    return w;
}

// G_m(r) for m < n: a bound on every order-m derivative of erfc(sqrt(p) R)/R
// for R >= r (src/electrostatic_potential.cpp:385-430):
//   G_m = (2/sqrt(pi)) k^3 / 2 sqrt(2^m m!) (2/r^2)^((m+1)/2) Gamma((m+1)/2, p r^2 / 2),
// k = 1.0865 (Cramer's bound on Hermite functions), the incomplete gamma
// values from Gamma(1/2, x) = sqrt(pi) erfc(sqrt(x)), Gamma(1, x) = e^-x and
// Gamma(s + 1, x) = s Gamma(s, x) + x^s e^-x. Infinite for r <= 0.
void erfc_tail_bounds(double p, double r, int n, double* out) {
    if (!(r > 0.0)) {
        std::fill(out, out + n, kInf);
        return;
    }
    const double x = 0.5 * p * r * r, ex = std::exp(-x), sx = std::sqrt(x);
    const double q = std::sqrt(2.0 / (r * r));
    constexpr double cramer = 1.0865;
    const double front = 2.0 / std::sqrt(std::numbers::pi) * cramer * cramer * cramer * 0.5; // (2/sqrt(pi)) k^3 / 2
    // Gamma((m+1)/2, x): even m from s = 1/2, odd m from s = 1
    double ge = std::sqrt(std::numbers::pi) * std::erfc(sx), go = ex;
    double pe = sx, po = x; // x^s of the next step of each sequence
    double se = 0.5, so = 1.0;
    double norm = 1.0, qpow = q;
    for (int m = 0; m < n; ++m) {
        if (m > 0) {
            norm *= 2.0 * m;
            qpow *= q;
        }
        double gamma;
        if (m % 2 == 0) {
            if (m > 0) {
                ge = se * ge + pe * ex;
                pe *= x;
                se += 1.0;
            }
            gamma = ge;
        } else {
            if (m > 1) {
                go = so * go + po * ex;
                po *= x;
                so += 1.0;
            }
            gamma = go;
        }
        out[m] = front * std::sqrt(norm) * qpow * gamma;
    }
}

// T_L(d) = sum_(n > K) (n + L)! / n! rho^n / d^(n + L + 1): the remainder of
// the order-K Taylor expansion of the order-L multipole terms over a ball of
// radius rho at distance d (src/electrostatic_potential.cpp:432-455); infinite
// unless d > 1.25 rho; the terms are summed until their ratio falls below
// 0.9 and the rest bounded by the geometric series.
double taylor_remainder(int L, int K, double rho, double d) {
    if (!(d > 1.25 * rho)) return kInf;
    const double ratio0 = rho / d;
    double term = 1.0 / std::pow(d, L + 1);
    for (int j = 2; j <= L; ++j)
        term *= j; // L! / d^(L+1)
    for (int n = 1; n <= K + 1; ++n)
        term *= (n + L) / static_cast<double>(n) * ratio0; // the n = K + 1 term
    double sum = 0.0;
    for (int n = K + 1; n < K + 1000; ++n) {
        sum += term;
        const double r = (n + L + 1) / static_cast<double>(n + 1) * ratio0;
        if (r < 0.9) return sum + term * r / (1.0 - r);
        term *= r;
        if (!std::isfinite(term)) break;
    }
    return kInf;
}

// The smallest d in [low, high] with bound(d) <= allowed, for a bound
// decreasing in d (src/electrostatic_potential.cpp:457-472): high doubled up
// to 1e4 bohr until it qualifies, then bisection to a relative 1e-3 (at most
// 40 halvings), the qualifying end returned.
template <typename Bound> double first_within(const Bound& bound, double low, double high, double allowed) {
    while (high < 1e4 && !(bound(high) <= allowed))
        high *= 2.0;
    if (!(bound(high) <= allowed)) return kInf;
    for (int it = 0; it < 40 && high - low > 1e-3 * high; ++it) {
        const double mid = 0.5 * (low + high);
        (bound(mid) <= allowed ? high : low) = mid;
    }
    return high;
}

// A double rounded up to the next float (src/electrostatic_potential.cpp:474-481).
float float_above(double v) {
    const float f = static_cast<float>(v);
    return static_cast<double>(f) < v ? std::nextafter(f, std::numeric_limits<float>::infinity()) : f;
}

} // namespace

FarPlanHost far_field_plan(const std::vector<EspPairDev>& pairs, const std::vector<double>& coef,
                           const std::vector<int32_t>& order, double radius, double tolerance) {
    if (!std::isfinite(radius) || radius < 0.0) invalid("cluster radius must be finite and nonnegative");
    if (!std::isfinite(tolerance) || !(tolerance > 0.0))
        invalid("the far-field plan needs a finite positive potential tolerance");
    FarPlanHost plan;
    const std::size_t np = pairs.size();
    plan.radius = radius;
    plan.tolerance = tolerance;
    plan.pair_tolerance = 0.5 * tolerance;
    for (int j = 0; j < kFarLevels; ++j)
        plan.levels[j] = plan.pair_tolerance * std::pow(10.0, -0.5 * j);
    plan.level_d2.assign(np * kFarLevels, std::numeric_limits<float>::infinity());
    for (const auto& pr : pairs)
        plan.max_L = std::max(plan.max_L, static_cast<int>(pr.L));
    plan.moment_count = far_hcount(plan.max_L + kFarMultipole);
    if (np == 0) return plan;
    if (order.size() != np) invalid("far-field plan: pair order does not match the pairs");

    // Error levels of each pair's local expansion at 64 distances from
    // max(1.25 rho, 0.05) to 400 bohr in geometric progression: the
    // truncation bound sum_L W_L [T_L(d) + G_L(d - rho) + sum_(n <= K)
    // (sqrt(3) rho)^n / n! G_(L+n)(d)], and for each level the first distance
    // whose bound is within it (src/electrostatic_potential.cpp:503-559).
    constexpr int kSamples = 64, K = kFarLocal;
    const double rho = radius;
    const double d_first = std::max(1.25 * rho, 0.05);
    const double step = std::pow(400.0 / d_first, 1.0 / (kSamples - 1));
    double dist[kSamples];
    for (int g = 0; g < kSamples; ++g)
        dist[g] = d_first * std::pow(step, g);
    std::vector<double> gnear, gfar;
    for (std::size_t i = 0; i < np; ++i) {
        const EspPairDev& pr = pairs[i];
        const std::vector<double> w = order_weights(pr, coef.data() + pr.first);
        gnear.assign(static_cast<std::size_t>(pr.L + 1), 0.0);
        gfar.assign(static_cast<std::size_t>(pr.L + K + 1), 0.0);
        float* lim = plan.level_d2.data() + i * kFarLevels;
        int j = 0;
        for (int g = 0; g < kSamples && j < kFarLevels; ++g) {
            const double d = dist[g];
            erfc_tail_bounds(pr.p, d - rho, pr.L + 1, gnear.data());
            erfc_tail_bounds(pr.p, d, pr.L + K + 1, gfar.data());
            double bound = 0.0;
            for (int L = 0; L <= pr.L; ++L) {
                const double wl = w[static_cast<std::size_t>(L)];
                if (wl == 0.0) continue;
                double term = taylor_remainder(L, K, rho, d) + gnear[static_cast<std::size_t>(L)];
                double scale = 1.0;
                for (int n = 0; n <= K; ++n) {
                    if (n > 0) scale *= std::sqrt(3.0) * rho / n;
                    term += scale * gfar[static_cast<std::size_t>(L + n)];
                }
                bound += wl * term;
            }
            while (j < kFarLevels && bound <= plan.levels[j])
                lim[j++] = float_above(d * d);
        }
    }

    // Source clusters: pairs keyed by the 2-bohr cube of their centre and the
    // exponent interval [4^k, 4^(k+1)), numbered by first appearance in the
    // reference's order and listing their pairs in it
    // (src/electrostatic_potential.cpp:561-699).
    std::map<std::array<long, 4>, int> key_index;
    std::vector<std::vector<int32_t>> members;
    plan.cluster_of.resize(np);
    for (const int32_t i : order) {
        const EspPairDev& pr = pairs[static_cast<std::size_t>(i)];
        const double q[3] = {pr.qx, pr.qy, pr.qz};
        std::array<long, 4> key{};
        for (int a = 0; a < 3; ++a)
            key[static_cast<std::size_t>(a)] = static_cast<long>(std::floor(q[a] / kFarCell));
        key[3] = static_cast<long>(std::floor(std::log2(pr.p) / 2.0));
        auto [it, fresh] = key_index.emplace(key, static_cast<int>(members.size()));
        if (fresh) members.emplace_back();
        members[static_cast<std::size_t>(it->second)].push_back(i);
        plan.cluster_of[static_cast<std::size_t>(i)] = it->second;
    }
    const int nclusters = static_cast<int>(members.size());
    constexpr int N = kFarMultipole;
    const int M = plan.moment_count;
    const int side = plan.max_L + N + 1;
    const std::vector<Tuv> all = tuv_list(plan.max_L + N);
    std::vector<int> pos(static_cast<std::size_t>(side * side * side), -1);
    for (int k = 0; k < M; ++k)
        pos[static_cast<std::size_t>((all[k].t * side + all[k].u) * side + all[k].v)] = k;
    const std::vector<Tuv> shifts = tuv_list(N);
    const double share = 0.25 * tolerance / nclusters;
    plan.clusters.resize(static_cast<std::size_t>(nclusters));
    plan.moments.assign(static_cast<std::size_t>(nclusters) * (N + 1) * M, 0.0);
    std::vector<double> by_order(static_cast<std::size_t>(N + 1) * M), tail;
    for (int ci = 0; ci < nclusters; ++ci) {
        const auto& list = members[static_cast<std::size_t>(ci)];
        FarClusterDev& cl = plan.clusters[static_cast<std::size_t>(ci)];
        cl.first = static_cast<int32_t>(plan.cluster_pairs.size());
        cl.count = static_cast<int32_t>(list.size());
        plan.cluster_pairs.insert(plan.cluster_pairs.end(), list.begin(), list.end());
        double lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
        double pmin = kInf;
        int top = 0;
        for (const int32_t i : list) {
            const EspPairDev& pr = pairs[static_cast<std::size_t>(i)];
            const double q[3] = {pr.qx, pr.qy, pr.qz};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], q[a]);
                hi[a] = std::max(hi[a], q[a]);
            }
            pmin = std::min(pmin, pr.p);
            top = std::max(top, static_cast<int>(pr.L));
        }
        for (int a = 0; a < 3; ++a)
            cl.c[a] = 0.5 * (lo[a] + hi[a]);
        // moments Q^(m)_nu = sum over pairs and terms alpha + delta = nu with
        // |delta| <= m of e_alpha u^delta / delta!, u = pair centre - cluster
        // centre, for m = 0..N (cumulative in m), and the weights per order
        double cradius = 0.0;
        std::vector<double> cw(static_cast<std::size_t>(top + 1), 0.0);
        std::fill(by_order.begin(), by_order.end(), 0.0);
        for (const int32_t i : list) {
            const EspPairDev& pr = pairs[static_cast<std::size_t>(i)];
            const double u[3] = {pr.qx - cl.c[0], pr.qy - cl.c[1], pr.qz - cl.c[2]};
            cradius = std::max(cradius, std::sqrt((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]));
            const std::vector<double> w = order_weights(pr, coef.data() + pr.first);
            for (std::size_t o = 0; o < w.size(); ++o)
                cw[o] += w[o];
            const double f = multipole_factor(pr.p);
            const double* c = coef.data() + pr.first;
            for (int k = 0; k < far_hcount(pr.L); ++k) {
                const double e = c[k] * f;
                if (e == 0.0) continue;
                for (const Tuv& d : shifts) {
                    double term = e;
                    for (int j = 1; j <= d.t; ++j)
                        term *= u[0] / j;
                    for (int j = 1; j <= d.u; ++j)
                        term *= u[1] / j;
                    for (int j = 1; j <= d.v; ++j)
                        term *= u[2] / j;
                    const int nu = pos[static_cast<std::size_t>(((all[k].t + d.t) * side + all[k].u + d.u) * side +
                                                                all[k].v + d.v)];
                    by_order[static_cast<std::size_t>(d.t + d.u + d.v) * M + nu] += term;
                }
            }
        }
        double* mom = plan.moments.data() + static_cast<std::size_t>(ci) * (N + 1) * M;
        for (int m = 0; m <= N; ++m)
            for (int nu = 0; nu < M; ++nu)
                mom[m * M + nu] = (m > 0 ? mom[(m - 1) * M + nu] : 0.0) + by_order[static_cast<std::size_t>(m) * M + nu];
        // translated only beyond both the Taylor truncation of its moments
        // over a target of the plan's radius and the erfc tail of its most
        // diffuse pair are within the share
        const double reach = cradius + radius;
        const auto truncation = [&](double D) {
            double t = 0.0;
            for (std::size_t o = 0; o < cw.size(); ++o)
                if (cw[o] > 0.0) t += cw[o] * taylor_remainder(static_cast<int>(o), N, reach, D);
            return t;
        };
        tail.assign(cw.size(), 0.0);
        const auto spread = [&](double r) {
            erfc_tail_bounds(pmin, r, static_cast<int>(tail.size()), tail.data());
            double t = 0.0;
            for (std::size_t o = 0; o < cw.size(); ++o)
                t += cw[o] * tail[o];
            return t;
        };
        const double dt = first_within(truncation, 1.25 * reach, 2.5 * reach + 1e-3, share);
        const double de = first_within(spread, 0.0, 1.0 / std::sqrt(pmin), share);
        cl.min_distance = std::max(dt, reach + de);
        cl.max_L = top;
    }
    return plan;
}

} // namespace kumo
