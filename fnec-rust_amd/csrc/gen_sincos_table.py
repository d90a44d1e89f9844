"""Writes sincos_table.h: for i = 0..110, sin(i/128) and cos(i/128), each as a
double-double pair (high part: the double nearest the value; low part: the
double nearest the remainder), from 300-bit mpmath evaluations. The table
serves the table-driven sin/cos of refmath.h.
Run from this directory:  python gen_sincos_table.py"""
import mpmath as mp

mp.mp.prec = 300
lines = [
    "// Generated by gen_sincos_table.py: for i = 0..110, sin(i/128), the remainder",
    "// sin(i/128) - that double, cos(i/128), and its remainder, each the nearest double",
    "// (300-bit evaluation). Used by the table-driven sin/cos of refmath.h.",
    "#pragma once",
    "#define FNEC_SINCOS_TABLE_LEN (111 * 4)",
    "#define FNEC_SINCOS_TABLE_INIT \\",
]
for i in range(111):
    x = mp.mpf(i) / 128
    s, c = mp.sin(x), mp.cos(x)
    sh, ch = float(s), float(c)
    vals = [sh, float(s - sh), ch, float(c - ch)]
    lines.append("    " + ", ".join(v.hex() if v != 0 else "0.0" for v in vals) + (", \\" if i < 110 else ""))
open("sincos_table.h", "w").write("\n".join(lines) + "\n")
