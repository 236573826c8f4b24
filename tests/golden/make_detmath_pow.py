"""Writes tests/golden/detmath_pow.json: known answers for the two correctly
rounded powers on the lighting path.

* ``reference``: the reference's own table of correctly rounded results
  (src/SourceSharp.Tests/MapFormats/Numerics/DetMathGoldens.cs, written by its
  tools/detmath_goldens.py from mpmath at 3000 bits): every row of kind F2
  ``Pow`` (DetMathF.Pow, float in and out) and S2 ``PowToSingle``
  (DetMath.PowToSingle, double in, rounded to double then to float), read as
  data.
* ``hard``: arguments whose double ``pow`` lands within 2^-49 of a float
  rounding boundary, found by the oracle's scanner over the ranges the path
  uses (the spot cone's multiplier in [2^-8, 1) under fringe exponents, some
  chosen where the plain (float)pow(double) rounds the wrong way, and two
  exact midpoints; the supersampling intensity I/256 under 1/2.2), with
  answers from mpmath at 200 bits.

Run from the repository root with the reference checked out at
/root/reference:  python tests/golden/make_detmath_pow.py
"""
from __future__ import annotations

import json
import os
import re
import struct
import sys

import mpmath
import numpy as np

REF = "/root/reference/src/SourceSharp.Tests/MapFormats/Numerics/DetMathGoldens.cs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "detmath_pow.json")


def f32(bits: int) -> float:
    return struct.unpack("<f", struct.pack("<I", bits))[0]


def f64(bits: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def bits32(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def reference_rows():
    src = open(REF).read()
    rows = re.findall(r'\("(F2|S2)", "(Pow|PowToSingle)", 0x([0-9A-F]+)UL, 0x([0-9A-F]+)UL, 0x([0-9A-F]+)UL\)',
                      src)
    out = []
    for kind, _fn, a, b, e in rows:
        out.append({"kind": kind, "a": int(a, 16), "b": int(b, 16), "expected": int(e, 16)})
    return out


def mp_round_f32(v) -> int:
    """The nearest float (ties to even) to the mpf v, as bits."""
    if v == 0:
        return 0
    neg = v < 0
    v = abs(v)
    e = int(mpmath.floor(mpmath.log(v, 2)))
    e = max(e, -126)
    q = v / mpmath.mpf(2) ** (e - 23)  # mantissa in units of the float ulp
    n = int(mpmath.floor(q))
    frac = q - n
    if frac > 0.5 or (frac == 0.5 and n % 2 == 1):
        n += 1
    val = float(mpmath.mpf(n) * mpmath.mpf(2) ** (e - 23))
    return bits32(-val if neg else val)


def mp_round_f64_then_f32(v) -> int:
    d = float(v)  # mpmath rounds to nearest double
    return bits32(float(np.float32(d)))


def hard_rows(oracle_lib):
    """Arguments that need the exact step, found by the oracle's scanner over
    consecutive floats, with answers from mpmath."""
    import ctypes

    scan = oracle_lib.vc_pow_scan
    scan.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_uint32, ctypes.c_int64, ctypes.c_int32,
                     ctypes.c_void_p, ctypes.c_int32]
    scan.restype = ctypes.c_int32
    mpmath.mp.prec = 200
    rows = []
    found = np.zeros(8, np.uint32)
    lo, one = bits32(2.0 ** -8), bits32(1.0)

    def add_f2(b, yf):
        v = mpmath.power(mpmath.mpf(f32(int(b))), mpmath.mpf(yf))
        rows.append({"kind": "F2", "a": int(b), "b": bits32(yf), "expected": mp_round_f32(v)})

    # spot multipliers in [2^-8, 1) under exponents of the fringe: the first
    # few arguments that take the exact step ...
    for y in (1.5, 2.5, 3.7, 0.7):
        yf = float(np.float32(y))
        n = scan(0, yf, lo, one - lo, 0, found.ctypes.data, 3)
        for b in found[:n]:
            add_f2(b, yf)
    # ... and arguments where (float)pow(double) rounds the wrong way
    misrounded = 0
    for y in (1.25, 1.5, 1.75, 2.25, 2.5, 2.75, 3.3, 3.7, 4.5, 6.1, 0.3, 0.6, 0.7, 0.9, 7.7, 9.5):
        yf = float(np.float32(y))
        n = scan(0, yf, lo, one - lo, 1, found.ctypes.data, 1)
        for b in found[:n]:
            add_f2(b, yf)
        misrounded += n
        if misrounded >= 2:
            break
    # exact midpoints: (b / 32)^5 for b = 29, 31 has 25 significant bits
    for b in (29, 31):
        x = b / 32.0
        rows.append({"kind": "F2", "a": bits32(x), "b": bits32(5.0),
                     "expected": mp_round_f32(mpmath.power(mpmath.mpf(x), 5))})
    # supersampling intensities: I / 256 under 1 / 2.2 for float I in [1, 2^16)
    y = 1.0 / 2.2
    n = scan(1, y, bits32(1.0), bits32(65536.0) - bits32(1.0), 0, found.ctypes.data, 6)
    for b in found[:n]:
        x = f32(int(b)) / 256.0
        v = mpmath.power(mpmath.mpf(x), mpmath.mpf(y))
        rows.append({"kind": "S2", "a": struct.unpack("<Q", struct.pack("<d", x))[0],
                     "b": struct.unpack("<Q", struct.pack("<d", y))[0],
                     "expected": mp_round_f64_then_f32(v)})
    return rows


def main() -> int:
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    import oracle

    L = oracle.lib()
    data = {"reference": reference_rows(), "hard": hard_rows(L),
            "source": "DetMathGoldens.cs F2 Pow / S2 PowToSingle rows; hard rows from mpmath at 200 bits"}
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=0)
    print(f"wrote {OUT}: {len(data['reference'])} reference rows, {len(data['hard'])} hard rows")
    return 0


if __name__ == "__main__":
    sys.exit(main())
