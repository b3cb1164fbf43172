"""Generates pow_tables.h, the lookup tables of the device pow (libm_pow.hpp).

The tables follow the published construction of the table-driven pow of
glibc >= 2.28 (ARM optimized-routines, Szabolcs Nagy 2018), which is what the
reference's Fortran ``**`` resolves to on the host:

  log part, N = 128 subintervals of z in [0x1.69555p-1, 0x1.69555p0) (bit
  pattern OFF = 0x3fe6955500000000 + i * 2^45), one entry per subinterval:
      1/c   = center < 1 ? round(N / center) / N : round(2N / center) / (2N)
      logc  = round(2^43 * log(c)) / 2^43
      tail  = (double)(log(c) - logc)
  so z/c - 1 is exact and k*ln2hi + logc has no rounding error;

  exp part, 2^(k/N) = H[k] (1 + T[k]) for k in [0, N):
      tab[2k]   = bits of T[k]
      tab[2k+1] = bits of H[k] - (k << 52) / N

Values are computed with 80-digit decimal arithmetic and rounded once.
Run:  python particles_amd/csrc/gen_pow_tables.py
"""
import os
import struct
from decimal import Decimal as D, getcontext

getcontext().prec = 80
N = 128
OFF = 0x3FE6955500000000


def bits_to_double(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def double_to_bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def log_table():
    rows = []
    for i in range(N):
        lo = D(bits_to_double(OFF + (i << 45)))
        hi = D(bits_to_double(OFF + ((i + 1) << 45)))
        center = (lo + hi) / 2
        if center < 1:
            invc = D(int((D(N) / center).to_integral_value())) / N
        else:
            invc = D(int((D(2 * N) / center).to_integral_value())) / (2 * N)
        logc_exact = -invc.ln()
        logc = (logc_exact * D(2) ** 43).to_integral_value() / D(2) ** 43
        rows.append((float(invc), float(logc), float(logc_exact - logc)))
    return rows


def exp_table():
    rows = []
    for k in range(N):
        v = D(2) ** (D(k) / N)
        h = float(v)
        t = float(v / D(h) - 1)
        rows.append((double_to_bits(t), (double_to_bits(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF))
    return rows


def main():
    out = ["// Generated by particles_amd/csrc/gen_pow_tables.py -- do not edit.",
           "// Tables of the table-driven pow (log part: N = 128 subintervals of",
           "// [0x1.69555p-1, 0x1.69555p0); exp part: 2^(k/128) = H[k] (1 + T[k])).",
           "#pragma once",
           "PTC_POW_TABLE const double ptc_pow_log_tab[128][3] = {"]
    out += ["    {%s, %s, %s}," % (a.hex(), b.hex(), c.hex()) for a, b, c in log_table()]
    out += ["};", "PTC_POW_TABLE const unsigned long long ptc_pow_exp_tab[256] = {"]
    out += ["    0x%016xULL, 0x%016xULL," % r for r in exp_table()]
    out += ["};"]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pow_tables.h")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
