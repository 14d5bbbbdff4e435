"""Generate the exact constants used by shp_trig.h.

pi is computed with Machin's formula in exact rational arithmetic; pi/2 is
split greedily into four non-overlapping doubles, and 1/n! is written as an
unevaluated double-double (hi + lo). Run: python gen_trig_constants.py
"""
from fractions import Fraction
from math import factorial


def arctan_inv(x, terms):
    # arctan(1/x) = sum (-1)^k / ((2k+1) x^(2k+1))
    s = Fraction(0)
    for k in range(terms):
        s += Fraction((-1) ** k, (2 * k + 1) * x ** (2 * k + 1))
    return s


def pi_fraction():
    # Machin: pi = 16 arctan(1/5) - 4 arctan(1/239); 80 terms give > 250 bits
    return 16 * arctan_inv(5, 80) - 4 * arctan_inv(239, 80)


def to_double(fr):
    return float(fr)  # Python rounds Fraction -> float correctly (round-half-even)


def split(fr, parts):
    out = []
    rem = fr
    for _ in range(parts):
        d = to_double(rem)
        out.append(d)
        rem -= Fraction(d)
    return out, rem


def main():
    pi = pi_fraction()
    pio2 = pi / 2
    parts, rem = split(pio2, 4)
    print("// pi/2 = P0 + P1 + P2 + P3 (+ %.3e)" % float(rem))
    for i, p in enumerate(parts):
        print("#define SHP_PIO2_%d %s" % (i, p.hex()))
    print("#define SHP_TWO_OVER_PI %s" % to_double(2 / pi).hex())
    print("// 1/n! as hi + lo")
    for n in range(2, 30):
        f = Fraction(1, factorial(n))
        hi = to_double(f)
        lo = to_double(f - Fraction(hi))
        print("  {%s, %s}, /* 1/%d! */" % (hi.hex(), lo.hex(), n))


if __name__ == "__main__":
    main()
