"""Writes phase_table.h: sin(j pi/32) and cos(j pi/32) for j = 0..63, each
the double nearest the value (300-bit mpmath evaluations). The table serves
the table-driven sin/cos of the fast fill and the far field (kernels.hip).
Run from this directory:  python gen_phase_table.py"""
import mpmath as mp

mp.mp.prec = 300
lines = [
    "// Generated by gen_phase_table.py: sin(j pi/32), cos(j pi/32) for j = 0..63, each",
    "// the nearest double (300-bit evaluation). Used by sincos_tab in kernels.hip.",
    "#pragma once",
    "#define FNEC_PHASE_TABLE_LEN 64",
    "#define FNEC_PHASE_TABLE_INIT \\",
]
for j in range(64):
    t = mp.mpf(j) / 32  # sinpi / cospi: exact zeros at the quarter turns
    vals = [float(mp.sinpi(t)), float(mp.cospi(t))]
    lines.append("    {" + ", ".join(v.hex() if v != 0 else "0.0" for v in vals) + "}" + (", \\" if j < 63 else ""))
open("phase_table.h", "w").write("\n".join(lines) + "\n")
