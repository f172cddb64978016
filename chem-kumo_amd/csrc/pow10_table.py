"""Writes pow10_table.hpp: 10^j = (hi + lo) * 2^sh for j in [JMIN, JMAX], with
hi in [1, 2) the double nearest to 10^j / 2^sh and lo the double nearest to
the remainder, so that hi + lo carries 10^j to a relative error below 2^-105.
The values are computed with exact rational arithmetic.

Usage: python3 pow10_table.py > pow10_table.hpp
"""
from fractions import Fraction

JMIN, JMAX = -310, 340


def nearest_double(q):
    """The double nearest to the rational q (ties to even): Python's int / int
    division, which Fraction.__float__ uses, rounds correctly."""
    return float(q)


def entry(j):
    v = Fraction(10) ** j
    sh = v.numerator.bit_length() - v.denominator.bit_length()
    if Fraction(2) ** sh > v:
        sh -= 1
    m = v / Fraction(2) ** sh
    assert 1 <= m < 2
    hi = nearest_double(m)
    assert 1.0 <= hi < 2.0
    lo = nearest_double(m - Fraction(hi))
    err = abs(Fraction(hi) + Fraction(lo) - m) / m
    assert err < Fraction(1, 2 ** 105), j
    return hi, lo, sh


def main():
    print("// Generated by pow10_table.py (exact rational arithmetic); do not edit.")
    print("// Entry j - kPow10Min: 10^j = (hi + lo) * 2^sh, hi in [1, 2), |hi + lo - 10^j 2^-sh| < 2^-105 hi.")
    print("#pragma once")
    print(f"#define KUMO_POW10_MIN ({JMIN})")
    print(f"#define KUMO_POW10_MAX ({JMAX})")
    print("#define KUMO_POW10_TABLE \\")
    rows = []
    for j in range(JMIN, JMAX + 1):
        hi, lo, sh = entry(j)
        rows.append(f"    {{{hi.hex()}, {lo.hex()}, {sh}}}")
    print(", \\\n".join(rows))


if __name__ == "__main__":
    main()
