"""Writes rt_exp_table.inc: the 2^(k/128) table of the table-driven exp in rt_libm.h.

For k = 0..127: H[k] = 2^(k/128) rounded to nearest double, T[k] = 2^(k/128)/H[k] - 1
rounded to nearest double; entry pair (bits(T[k]), bits(H[k]) - (k << 45)).  Values are
computed with 80-digit decimal arithmetic.
"""
import os
import struct
from decimal import Decimal, getcontext

getcontext().prec = 80
N = 128
ln2 = Decimal(2).ln()
lines = []
for k in range(N):
    v = (ln2 * k / N).exp()
    h = float(v)
    t = float(v / Decimal(h) - 1)
    hb = struct.unpack("<Q", struct.pack("<d", h))[0]
    tb = struct.unpack("<Q", struct.pack("<d", t))[0]
    lines.append(f"0x{tb:016x}ULL, 0x{(hb - (k << 45)) & 0xffffffffffffffff:016x}ULL,")
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rt_exp_table.inc"), "w") as f:
    f.write("// generated by gen_exp_table.py: {bits(T[k]), bits(H[k]) - (k << 45)}, k = 0..127\n")
    f.write("\n".join(lines) + "\n")
