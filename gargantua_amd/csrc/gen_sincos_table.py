"""Writes gg_sincos_table.inc: the constants of sincos_cr (gg_crtrig.hpp).

sin(j/256) and cos(j/256) for j = 0..201 (0 .. pi/4 + 1/512) as double-double
pairs (hi = the double nearest the value, lo = the double nearest the rest),
and pi/2 in three parts, 2/pi, from 60-digit mpmath values.  Run from any
directory: python gargantua_amd/csrc/gen_sincos_table.py
"""
import os

import mpmath

mpmath.mp.dps = 60


def dd(v):
    hi = float(v)
    lo = float(v - mpmath.mpf(hi))
    return hi, lo


def main():
    rows = []
    for j in range(202):
        a = mpmath.mpf(j) / 256
        s_hi, s_lo = dd(mpmath.sin(a))
        c_hi, c_lo = dd(mpmath.cos(a))
        rows.append(f"    {{{s_hi!r}, {s_lo!r}, {c_hi!r}, {c_lo!r}}},")
    half_pi = mpmath.pi / 2
    p1 = float(half_pi)
    p2 = float(half_pi - mpmath.mpf(p1))
    p3 = float(half_pi - mpmath.mpf(p1) - mpmath.mpf(p2))
    out = [
        "// generated by gen_sincos_table.py: do not edit",
        "// sin(j/256), cos(j/256) for j = 0..201 as (sin hi, sin lo, cos hi, cos lo)",
        "GG_SINCOS_CONST double kSinCosTab[202][4] = {",
        *rows,
        "};",
        f"constexpr double kPio2_1 = {p1!r};   // pi/2 = kPio2_1 + kPio2_2 + kPio2_3 + O(1e-50)",
        f"constexpr double kPio2_2 = {p2!r};",
        f"constexpr double kPio2_3 = {p3!r};",
        f"constexpr double kTwoOverPi = {float(2 / mpmath.pi)!r};",
        "",
    ]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gg_sincos_table.inc")
    with open(path, "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
