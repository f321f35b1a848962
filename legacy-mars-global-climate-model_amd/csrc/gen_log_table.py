"""Writes rt_log_table.inc: log(c) for the table-driven correctly rounded log in rt_libm.h.

Entries 0..52: c = 1 + j/128 (j = 0..52), the table points of m in [1, 1.4140625);
entries 53..128: c = 1 - j/256 (j = 0..75), those of m in [0.70703125, 1).  Each log(c)
is evaluated with 60-digit decimal arithmetic and split into a double-double
{hi, lo}: hi = log(c) rounded to nearest, lo = (log(c) - hi) rounded to nearest.
"""
import os
import struct
from decimal import Decimal, getcontext

getcontext().prec = 60


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


cs = [Decimal(1) + Decimal(j) / 128 for j in range(53)] + [Decimal(1) - Decimal(j) / 256 for j in range(76)]
lines = []
for c in cs:
    v = c.ln()
    hi = float(v)
    lo = float(v - Decimal(hi))
    lines.append(f"0x{bits(hi):016x}ULL, 0x{bits(lo):016x}ULL,")
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rt_log_table.inc"), "w") as f:
    f.write("// generated by gen_log_table.py: {bits(log(c) hi), bits(log(c) lo)}, "
            "c = 1 + j/128 (j = 0..52), then c = 1 - j/256 (j = 0..75)\n")
    f.write("\n".join(lines) + "\n")
