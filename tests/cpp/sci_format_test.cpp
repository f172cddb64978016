// The exact decimal formatter of the Cube writer (chem-kumo_amd/csrc/
// sci_format.hpp), compiled for the host, against glibc's snprintf("%.*e")
// and snprintf("%13.5e") (the text libstdc++ writes for the reference's
// std::scientific / setprecision(5) / setw(13), src/cube.cpp:104-109).
// Inputs: random bit patterns over the whole double range (subnormals
// included), exact decimal ties (rounded to even) and their neighbours,
// powers of ten and their neighbours, values that round up to the next
// power of ten, and the extremes. Each is checked on the fast path and,
// separately, with every rounding decided on the exact big-integer path.
// Usage: sci_format_test [random count]; exit status 0 when all match.
#include "../../chem-kumo_amd/csrc/sci_format.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

const kumo::sci::Pow10 kTable[] = {KUMO_POW10_TABLE};

long long g_checked = 0, g_failed = 0;

void check(double v, int P, bool exact) {
    char want[64], got[64];
    std::snprintf(want, sizeof want, "%.*e", P, v);
    const int n = kumo::sci::format_sci(v, P, kTable, exact, got);
    got[n] = '\0';
    ++g_checked;
    if (std::strcmp(want, got) != 0) {
        if (++g_failed <= 20)
            std::printf("mismatch P=%d exact=%d v=%a: want %s got %s\n", P, exact ? 1 : 0, v, want, got);
        return;
    }
    if (P == 5) {
        char w13[32], g13[14];
        std::snprintf(w13, sizeof w13, "%13.5e", v);
        kumo::sci::format_cube_sample(v, kTable, exact, g13);
        if (std::strlen(w13) != 13 || std::memcmp(w13, g13, 13) != 0) {
            if (++g_failed <= 20) std::printf("width-13 mismatch v=%a: want [%s]\n", v, w13);
        }
    }
}

void check_all(double v) {
    for (int P : {0, 1, 5, 8, 12})
        for (bool exact : {false, true}) {
            check(v, P, exact);
            check(-v, P, exact);
        }
}

// Exact decimal ties of P + 1 = 6 significant digits: (2f + 1) 10^k / 2 with
// an odd part small enough to be a double.
void ties(std::mt19937_64& rng, std::vector<double>& out) {
    for (int k = -9; k <= 14; ++k) {
        for (int trial = 0; trial < 400; ++trial) {
            const uint64_t f = 100000 + rng() % 900000;
            uint64_t odd = 2 * f + 1;
            double v;
            if (k >= 0) {
                uint64_t p5 = 1;
                bool fits = true;
                for (int i = 0; i < k; ++i) {
                    if (p5 > (uint64_t(1) << 53) / 5) fits = false;
                    p5 *= 5;
                }
                if (!fits || odd > (uint64_t(1) << 53) / p5) continue;
                v = std::ldexp(static_cast<double>(odd * p5), k - 1);
            } else {
                uint64_t p5 = 1;
                for (int i = 0; i < -k; ++i) p5 *= 5;
                const uint64_t q = odd / p5;
                if (q == 0) continue;
                odd = (q | 1) * p5; // an odd multiple of 5^-k
                if (odd < 200001 || odd > 1999999) continue;
                v = std::ldexp(static_cast<double>(odd / p5), k - 1);
            }
            out.push_back(v);
            out.push_back(std::nextafter(v, 0.0));
            out.push_back(std::nextafter(v, INFINITY));
        }
    }
}

} // namespace

int main(int argc, char** argv) {
    const long long nrandom = argc > 1 ? std::atoll(argv[1]) : 300000;
    std::mt19937_64 rng(20261203);
    std::vector<double> special = {0.0,
                                   4.9406564584124654e-324,
                                   2.2250738585072009e-308,
                                   2.2250738585072014e-308,
                                   1.7976931348623157e308,
                                   1.0,
                                   0.5,
                                   0.1,
                                   1.25,
                                   9.999995,
                                   9.9999949999999999,
                                   999999.5,
                                   1e22,
                                   1e23,
                                   5e-324 * 3};
    for (int e = -323; e <= 308; ++e) {
        const double p = std::strtod(("1e" + std::to_string(e)).c_str(), nullptr);
        for (double v : {p, std::nextafter(p, 0.0), std::nextafter(p, INFINITY), 9.999995 * p, 9.9999951 * p,
                         9.9999949 * p, 5.0000005 * p})
            if (std::isfinite(v) && v != 0.0) special.push_back(v);
    }
    ties(rng, special);
    for (double v : special) check_all(v);
    const long long special_checked = g_checked;
    for (long long i = 0; i < nrandom; ++i) {
        uint64_t bits = rng() & 0x7fffffffffffffffull;
        if ((bits >> 52) == 0x7ff) continue; // infinities and NaNs are rejected before formatting
        double v;
        std::memcpy(&v, &bits, sizeof v);
        for (int P : {5, 12}) {
            check(v, P, false);
            check(-v, P, false);
        }
        if (i % 16 == 0) {
            check(v, 5, true);
            check(v, 12, true);
        }
    }
    // Densities and amplitudes as a field holds them: log-uniform over 60 decades.
    std::uniform_real_distribution<double> u(-50.0, 10.0);
    for (long long i = 0; i < nrandom; ++i) check(std::pow(10.0, u(rng)) * (rng() & 1 ? 1 : -1), 5, false);
    std::printf("%lld formats checked (%lld special), %lld mismatches\n", g_checked, special_checked, g_failed);
    return g_failed == 0 ? 0 : 1;
}
