"""Writes eco_pow_tables.h, the two 128-entry tables of the isentrope's power
(eco_math.h pow_pos), from their defining formulas in 300-bit arithmetic
(mpmath).  The tables are those of the table-driven pow published by
Szabolcs Nagy in ARM's optimized-routines (math/pow.c, pow_log_data.c,
exp_data.c), the pow of the GNU C library since 2.28:

log table, N = 128 subintervals of z in [0x1.69555p-1, 0x1.69555p0) (bit
patterns OFF + i 2^45 .. OFF + (i+1) 2^45, OFF = 0x3fe6955500000000), c near
the subinterval's centre with 1/c = round(N/centre)/N (centre < 1) or
round(2N/centre)/2N, so that z/c - 1 is exact in one fma:
    invc = 1/c, logc = round(2^43 log c)/2^43, logctail = RN(log c - logc)

exp table, 2^(k/N) ~= H_k (1 + T_k), k = 0..N-1:
    word 2k = bits of T_k = RN(2^(k/N)/H_k - 1), H_k = RN(2^(k/N))
    word 2k+1 = bits of H_k - (k << 45)

    python ecogen_amd/csrc/gen_pow_tables.py    # rewrites eco_pow_tables.h
"""
import struct
from pathlib import Path

import mpmath as mp

mp.mp.prec = 300
N = 128
OFF = 0x3FE6955500000000


def _asd(u: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", u & 0xFFFFFFFFFFFFFFFF))[0]


def _asu(d: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def log_table():
    rows = []
    for i in range(N):
        lo, hi = _asd(OFF + (i << 45)), _asd(OFF + ((i + 1) << 45))
        centre = (mp.mpf(lo) + mp.mpf(hi)) / 2
        invc = mp.nint(N / centre) / N if centre < 1 else mp.nint(2 * N / centre) / (2 * N)
        invc_d = float(invc)
        assert mp.mpf(invc_d) == invc
        logc_exact = -mp.log(mp.mpf(invc_d))
        logc = mp.nint(mp.mpf(2) ** 43 * logc_exact) / mp.mpf(2) ** 43
        logc_d = float(logc)
        assert mp.mpf(logc_d) == logc
        rows.append((invc_d, logc_d, float(logc_exact - logc)))
    return rows


def exp_table():
    words = []
    for k in range(N):
        v = mp.mpf(2) ** (mp.mpf(k) / N)
        h = float(v)
        t = float(v / mp.mpf(h) - 1)
        words += [_asu(t), (_asu(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF]
    return words


def main() -> None:
    out = Path(__file__).resolve().parent / "eco_pow_tables.h"
    lines = ["// Generated by gen_pow_tables.py (do not edit): the log and exp tables of",
             "// pow_pos in eco_math.h.  Rows of ECO_POW_LOG_ROWS: {1/c, logc, logctail};",
             "// words of ECO_POW_EXP_WORDS: T_k, H_k - (k << 45) for 2^(k/128) ~= H_k (1 + T_k).",
             "#pragma once", "", "#define ECO_POW_LOG_ROWS \\"]
    lines += [f"  {{{a.hex()}, {b.hex()}, {c.hex()}}}, \\" for a, b, c in log_table()]
    lines += ["", "#define ECO_POW_EXP_WORDS \\"]
    w = exp_table()
    lines += [f"  0x{w[2 * k]:016x}ull, 0x{w[2 * k + 1]:016x}ull, \\" for k in range(N)]
    lines += [""]
    out.write_text("\n".join(lines) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
