"""Print the constants of cl_math.h from their mathematical definitions (mpmath, 1800 bits).

Run: python chaos_amd/csrc/gen_math_constants.py > /tmp/consts.txt, then compare with the header.
Every constant is a rounding of an exact value: Taylor coefficients 1/n!, the atanh series
coefficients 2/(2i+1), pi/2 split into doubles, ln 2 split with a short high part, and the
binary digits of 2/pi for the Payne-Hanek reduction.
"""
import mpmath as mp
from fractions import Fraction
from math import factorial

mp.mp.prec = 1800


def hexd(v):
    return float(v).hex()


def split(value, parts, hi_bits=None):
    out = []
    rest = mp.mpf(value)
    for i in range(parts):
        if i == 0 and hi_bits is not None:
            m, e = mp.frexp(rest)
            d = mp.floor(m * 2**hi_bits) / 2**hi_bits * mp.mpf(2) ** e
        else:
            d = mp.mpf(float(rest))
        out.append(float(d))
        rest = rest - mp.mpf(float(d))
    return out


if __name__ == "__main__":
    pio2 = mp.pi / 2
    print("// pi/2 = C1 + C2 + C3 + C4")
    for i, c in enumerate(split(pio2, 4)):
        print(f"#define CL_PIO2_{i+1} {hexd(c)}")
    print(f"#define CL_INV_PIO2 {hexd(2 / mp.pi)}")
    print("// ln 2 = LN2_HI (42 bits) + LN2_LO")
    hi, lo = split(mp.log(2), 2, hi_bits=42)
    print(f"#define CL_LN2_HI {hexd(hi)}\n#define CL_LN2_LO {hexd(lo)}")
    print("// sin: (-1)^i / (2i+1)!, i = 1..10")
    for i in range(1, 11):
        print(f"  {hexd(Fraction((-1) ** i, factorial(2 * i + 1)))},  // {'-' if i % 2 else '+'}1/{2*i+1}!")
    print("// cos: (-1)^i / (2i)!, i = 2..11")
    for i in range(2, 12):
        print(f"  {hexd(Fraction((-1) ** i, factorial(2 * i)))},  // {'-' if i % 2 else '+'}1/{2*i}!")
    print("// log: 2/(2i+1), i = 1..12")
    for i in range(1, 13):
        print(f"  {hexd(Fraction(2, 2 * i + 1))},  // 2/{2*i+1}")
    print("// bits of 2/pi, 64 per word, most significant first (word j holds bits 64j+1 .. 64j+64)")
    v = 2 / mp.pi
    words = []
    for j in range(24):
        v = v * 2**64
        w = int(mp.floor(v))
        words.append(w)
        v = v - w
    for j in range(0, 24, 3):
        print("  " + ", ".join(f"0x{w:016x}ull" for w in words[j:j + 3]) + ",")
