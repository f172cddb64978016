// Exact decimal formatting of doubles in the text of printf("%.*e", P, v)
// (glibc, round to nearest, ties to even), for host and device code.
//
// Field::write_cube formats every sample as iostream's std::scientific with
// precision 5 in a field of width 13 (src/cube.cpp:104-109), which libstdc++
// produces with vsnprintf("%.*e"). The digits here are those of the exact
// binary value of v rounded to P + 1 significant digits:
//  - fast path: y = v 10^(P - e10) in double-double arithmetic, |error| below
//    2^-100 y, from the exact table of powers of ten in pow10_table.hpp; the
//    rounding is decided by the sign of y - (floor(y) + 1/2) wherever that
//    difference exceeds kTieMargin in magnitude (far above the error);
//  - exact path (|y - floor(y) - 1/2| <= kTieMargin: ties and near ties):
//    compares 2 M 2^E 10^j with 2 floor(y) + 1 in big-integer arithmetic
//    (v = M 2^E, j = P - e10) and rounds ties to even.
// Contraction of a * b + c into fma must be off for this arithmetic (a
// clang pragma in sci_digits; compile host code with -ffp-contract=off).
// Every finite double formats to between P + 6 and P + 8 characters ("-",
// "d.", P digits, "e", sign, 2 or 3 exponent digits), so with P = 5 the
// width-13 field is never exceeded and each sample takes exactly 13 bytes.
#pragma once

#include "pow10_table.hpp"

#include <cstdint>

#if defined(__HIPCC__)
#define KUMO_HD __host__ __device__
#else
#define KUMO_HD
#endif

namespace kumo::sci {

struct Pow10 {
    double hi, lo;
    int32_t sh;
};
constexpr int kPow10Min = KUMO_POW10_MIN;
constexpr int kPow10Max = KUMO_POW10_MAX;
constexpr int kMaxPrecision = 12; // digits after the point supported (10^(P+1) < 2^53)
constexpr double kTieMargin = 1e-13;

// Fixed-size unsigned big integer, 32-bit limbs, least significant first.
// 36 limbs (1152 bits) hold every operand of the exact comparison: both
// sides are about 2 y 2^max(-E-j, 0) 5^max(-j, 0) with y < 10^13, below
// 2^810 over the whole double range.
struct BigInt {
    static constexpr int kLimbs = 36;
    uint32_t w[kLimbs];
    int n;
    KUMO_HD void set(uint64_t v) {
        n = 0;
        while (v) {
            w[n++] = static_cast<uint32_t>(v);
            v >>= 32;
        }
    }
    KUMO_HD void mul(uint32_t m) {
        uint64_t carry = 0;
        for (int i = 0; i < n; ++i) {
            const uint64_t t = static_cast<uint64_t>(w[i]) * m + carry;
            w[i] = static_cast<uint32_t>(t);
            carry = t >> 32;
        }
        if (carry) w[n++] = static_cast<uint32_t>(carry);
    }
    KUMO_HD void mul_pow5(int e) {
        while (e >= 13) {
            mul(1220703125u); // 5^13
            e -= 13;
        }
        uint32_t m = 1;
        while (e-- > 0) m *= 5;
        if (m != 1) mul(m);
    }
    KUMO_HD void shl(int s) {
        if (n == 0) return;
        const int limbs = s >> 5, bits = s & 31;
        if (bits) {
            uint32_t carry = 0;
            for (int i = 0; i < n; ++i) {
                const uint32_t v = w[i];
                w[i] = (v << bits) | carry;
                carry = v >> (32 - bits);
            }
            if (carry) w[n++] = carry;
        }
        if (limbs) {
            for (int i = n - 1; i >= 0; --i) w[i + limbs] = w[i];
            for (int i = 0; i < limbs; ++i) w[i] = 0;
            n += limbs;
        }
    }
};

KUMO_HD inline int compare(const BigInt& a, const BigInt& b) {
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}

// Sign of M 2^E 10^j - (f + 1/2), exactly: compares 2 M 2^(E + j) 5^j with
// 2 f + 1, negative powers moved to the other side. Inlined like every device
// function here (its big integers then live in the kernel's own private
// segment; the path is taken only for ties and near ties).
__attribute__((always_inline)) KUMO_HD inline int half_compare(uint64_t M, int E, int j, uint64_t f) {
    BigInt a, b;
    a.set(2 * M);
    b.set(2 * f + 1);
    if (j >= 0)
        a.mul_pow5(j);
    else
        b.mul_pow5(-j);
    const int p = E + j;
    if (p >= 0)
        a.shl(p);
    else
        b.shl(-p);
    return compare(a, b);
}

struct Digits {
    uint64_t d; // P + 1 significant digits, 10^P <= d < 10^(P+1) (0 for zero)
    int e10;    // decimal exponent of the first digit
};

KUMO_HD inline uint64_t pow10u(int e) {
    uint64_t r = 1;
    while (e-- > 0) r *= 10;
    return r;
}

// Significant digits of |v| rounded to P + 1 digits (v finite; 0 <= P <=
// kMaxPrecision). exact_only: decide every rounding on the exact path (for
// testing that path).
KUMO_HD inline Digits sci_digits(double v, int P, const Pow10* table, bool exact_only) {
#if defined(__clang__)
    // The error-free transformations below need every product and sum
    // rounded on its own: a multiply fused into a later add (which the AMDGPU
    // backend does even when the product has other uses) loses the low part.
#pragma clang fp contract(off)
#endif
    const uint64_t bits = __builtin_bit_cast(uint64_t, v) & 0x7fffffffffffffffull;
    if (bits == 0) return {0, 0};
    const int be = static_cast<int>(bits >> 52);
    uint64_t M = bits & 0xfffffffffffffull;
    int E;
    if (be == 0) {
        E = -1074;
    } else {
        M |= 1ull << 52;
        E = be - 1075;
    }
    // floor(log10 2^t) <= e10 <= floor(log10 2^t) + 1 for t = floor(log2 |v|);
    // t log10(2) is at least 4e-4 from an integer for 0 < |t| < 2136.
    const int t = E + 63 - __builtin_clzll(M);
    int e10 = static_cast<int>(__builtin_floor(t * 0.30102999566398120));
    const uint64_t top = pow10u(P + 1);
    const double topd = static_cast<double>(top);
    const double Md = static_cast<double>(M);
    for (int pass = 0;; ++pass) {
        const int j = P - e10;
        const Pow10& p = table[j - kPow10Min];
        const double ph = Md * p.hi;
        const double pl = __builtin_fma(Md, p.hi, -ph);
        const double s = __builtin_fma(Md, p.lo, pl);
        double yh = ph + s;
        double yl = s - (yh - ph);
        yh = __builtin_ldexp(yh, E + p.sh);
        yl = __builtin_ldexp(yl, E + p.sh);
        // The estimate of e10 was low: y >= 10^(P+1). Near that boundary both
        // choices round to 10^(P+1) (then carried to 10^P, e10 + 1).
        if (pass == 0 && yh >= topd) {
            ++e10;
            continue;
        }
        const double fl = __builtin_floor(yh);
        const double g = (yh - fl) - 0.5; // exact: yh and fl share a binade
        const double dist = g + yl;
        const uint64_t f = static_cast<uint64_t>(fl);
        uint64_t d;
        if (exact_only || __builtin_fabs(dist) <= kTieMargin) {
            const int c = half_compare(M, E, j, f);
            d = c > 0 ? f + 1 : (c < 0 ? f : f + (f & 1));
        } else {
            d = dist > 0 ? f + 1 : f;
        }
        if (d >= top) {
            d /= 10; // 10^(P+1) -> 10^P
            ++e10;
        }
        return {d, e10};
    }
}

// printf("%.*e", P, v) text of a finite v into out (no terminator); returns
// its length.
KUMO_HD inline int format_sci(double v, int P, const Pow10* table, bool exact_only, char* out) {
    const Digits dg = sci_digits(v, P, table, exact_only);
    int n = 0;
    if (__builtin_bit_cast(uint64_t, v) >> 63) out[n++] = '-';
    char digits[kMaxPrecision + 1] = {};
    uint64_t d = dg.d;
    for (int i = P; i >= 0; --i) {
        digits[i] = static_cast<char>('0' + d % 10);
        d /= 10;
    }
    out[n++] = digits[0];
    if (P > 0) {
        out[n++] = '.';
        for (int i = 1; i <= P; ++i) out[n++] = digits[i];
    }
    out[n++] = 'e';
    int e = dg.e10;
    out[n++] = e < 0 ? '-' : '+';
    if (e < 0) e = -e;
    if (e >= 100) {
        out[n++] = static_cast<char>('0' + e / 100);
        e %= 100;
    }
    out[n++] = static_cast<char>('0' + e / 10);
    out[n++] = static_cast<char>('0' + e % 10);
    return n;
}

// The 13 bytes of printf("%13.5e", v) for a finite v.
KUMO_HD inline void format_cube_sample(double v, const Pow10* table, bool exact_only, char* out13) {
    char s[16];
    const int n = format_sci(v, 5, table, exact_only, s);
    const int pad = 13 - n;
    for (int i = 0; i < pad; ++i) out13[i] = ' ';
    for (int i = 0; i < n; ++i) out13[pad + i] = s[i];
}

} // namespace kumo::sci
