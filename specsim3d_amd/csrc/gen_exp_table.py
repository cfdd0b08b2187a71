"""Write ss3d_exp_table.inc: the 128-entry table of 2^(k/128) used by ss3d_exp.h.

Entry pair k: (tail bits, scale bits - (k << 52)/128), where H = RN(2^(k/128)) is the scale and
tail = RN(2^(k/128)/H - 1), so that 2^(k/128) ~= H (1 + tail).  Computed with mpmath at 300 bits.

Usage:  python specsim3d_amd/csrc/gen_exp_table.py   (mpmath; run in the build container)
"""
import os
import struct

import mpmath

N = 128


def table():
    mpmath.mp.prec = 300
    bits = lambda x: struct.unpack("<Q", struct.pack("<d", x))[0]
    out = []
    for k in range(N):
        e = mpmath.power(2, mpmath.mpf(k) / N)
        h = float(e)  # round to nearest
        t = float(e / mpmath.mpf(h) - 1)
        out.append((bits(t), (bits(h) - ((k << 52) // N)) & 0xFFFFFFFFFFFFFFFF))
    return out


def text():
    lines = ["// Generated by gen_exp_table.py: pairs (tail, scale - (k << 52)/128) of 2^(k/128), k = 0..127."]
    for t, h in table():
        lines.append("    0x%016xull, 0x%016xull," % (t, h))
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ss3d_exp_table.inc")
    with open(path, "w") as fh:
        fh.write(text())
    print(path)
