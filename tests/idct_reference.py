"""The reference's f64 IDCTs evaluated under IEEE-754 rules, for the oracle's exactness pins
(tests/test_oracle_idct_pins.py) and their fixture generator (tests/golden/make_idct_pins.py).

Loop orders, from the reference:
  inverse_dct (MJPEG crates/codecs/mjpeg/src/transform.rs:86-109, MPEG-2
  crates/codecs/mpeg2/src/transform.rs:90-110): tmp[v*8+x] += M[x][u] * c[v*8+u] (u ascending,
  from 0.0); sum += M[y][v] * tmp[v*8+x] (v ascending, from 0.0); MJPEG then (sum + 128.0).round()
  clamped to 0..255, MPEG-2 sum.round() clamped to -256..255;
  inverse_88 (crates/codecs/dv/src/decoder.rs:520-548): for (y, x), v outer, u inner, sum +=
  cu * cv * c * cos16[x][u] * cos16[y][v] (left to right), then (sum / 4.0).round() clamped 0..255;
  inverse_248 (:555-602): row-pair sums / differences, horizontal (sum of cu * b * cos16[x][u]) / 2,
  per field vertical (sum of cv * h * cos8[y][v]) / 2, round, clamp 0..255.
Three evaluations: `plain` with Python floats (binary64, round to nearest even after every
operation, never fused: exactly the reference's arithmetic); `fused` with exact rationals rounded
once per a * b + s (what FMA contraction computes); `exact`, the real-number value of the same sum
over the same binary64 constants (no rounding at all), with a bound on how far any binary64
evaluation of it can be.
The basis constants are oracle/tables.h's ORACLE_IDCT_MATRIX; the DV cosine tables are the ones the
oracle computes with the C library's cos (the reference's f64::cos on Linux calls the same cos).
"""
import ctypes as C
import math
import os
import re
from fractions import Fraction as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SQRT1_2 = float.fromhex("0x1.6a09e667f3bcdp-1")   # FRAC_1_SQRT_2 as a binary64
U = 2.0 ** -53                                     # unit roundoff of binary64


def basis():
    src = open(os.path.join(ROOT, "oracle", "tables.h")).read()
    body = src[src.index("ORACLE_IDCT_MATRIX[8][8] = {"):]
    body = body[body.index("{") + 1:body.index("};")]
    vals = [float(v) for v in re.findall(r"-?\d+\.\d+(?:[eE]-?\d+)?", body)]
    assert len(vals) == 64
    return [vals[i * 8:(i + 1) * 8] for i in range(8)]


def dv_tables(lib):
    lib.oracle_dv_cos16_table.restype = C.POINTER(C.c_double)
    lib.oracle_dv_cos8_table.restype = C.POINTER(C.c_double)
    c16, c8 = lib.oracle_dv_cos16_table(), lib.oracle_dv_cos8_table()
    return [c16[i] for i in range(64)], [c8[i] for i in range(32)]


def round_half_away(q):
    """Rust's f64::round on an exact value (float or Fraction): half-way cases away from zero."""
    q = F(q)
    return int(math.copysign(math.floor(abs(q) + F(1, 2)), q))


def clamp(v, lo, hi):
    return lo if v < lo else hi if v > hi else v


class Transforms:
    """The four IDCTs in the three evaluations; `kind` is one of KINDS."""

    KINDS = ("mjpeg", "mpeg2", "dv88", "dv248")

    def __init__(self, lib):
        self.M = basis()
        self.c16, self.c8 = dv_tables(lib)
        self.MF = [[F(v) for v in r] for r in self.M]
        self.c16F, self.c8F = [F(v) for v in self.c16], [F(v) for v in self.c8]

    # ---- final rounding of each transform (on the pre-round value)
    @staticmethod
    def finish(kind, v):
        if kind == "mpeg2":
            return clamp(round_half_away(v), -256, 255)
        return clamp(round_half_away(v), 0, 255)

    # ---- plain binary64 (the reference)
    def plain(self, kind, c):
        return getattr(self, "_plain_" + kind)(c)

    def _separable_plain(self, c):
        M = self.M
        tmp = [0.0] * 64
        for v in range(8):
            for x in range(8):
                s = 0.0
                row = M[x]
                for u in range(8):
                    s = s + row[u] * c[v * 8 + u]
                tmp[v * 8 + x] = s
        out = [0.0] * 64
        for y in range(8):
            row = M[y]
            for x in range(8):
                s = 0.0
                for v in range(8):
                    s = s + row[v] * tmp[v * 8 + x]
                out[y * 8 + x] = s
        return out

    def _plain_mjpeg(self, c):
        return [s + 128.0 for s in self._separable_plain(c)]

    def _plain_mpeg2(self, c):
        return self._separable_plain(c)

    def _plain_dv88(self, c):
        c16 = self.c16
        out = [0.0] * 64
        for y in range(8):
            for x in range(8):
                s = 0.0
                for v in range(8):
                    cv = SQRT1_2 if v == 0 else 1.0
                    for u in range(8):
                        cu = SQRT1_2 if u == 0 else 1.0
                        t = cu * cv
                        t = t * c[v * 8 + u]
                        t = t * c16[x * 8 + u]
                        t = t * c16[y * 8 + v]
                        s = s + t
                out[y * 8 + x] = s / 4.0
        return out

    def _plain_dv248(self, c):
        c16, c8 = self.c16, self.c8
        b = [0.0] * 64
        for pair in range(4):
            for u in range(8):
                first, second = float(c[pair * 16 + u]), float(c[pair * 16 + 8 + u])
                b[pair * 16 + u] = first + second
                b[pair * 16 + 8 + u] = first - second
        h = [0.0] * 64
        for row in range(8):
            for x in range(8):
                s = 0.0
                for u in range(8):
                    t = (SQRT1_2 if u == 0 else 1.0) * b[row * 8 + u]
                    t = t * c16[x * 8 + u]
                    s = s + t
                h[row * 8 + x] = s / 2.0
        out = [0.0] * 64
        for field in range(2):
            for y in range(4):
                for x in range(8):
                    s = 0.0
                    for v in range(4):
                        t = (SQRT1_2 if v == 0 else 1.0) * h[(v * 2 + field) * 8 + x]
                        t = t * c8[y * 4 + v]
                        s = s + t
                    out[(y * 2 + field) * 8 + x] = s / 2.0
        return out

    # ---- fused: each accumulation step a * b + s rounded once
    def fused(self, kind, c):
        r = lambda q: F(float(q))   # nearest binary64, ties to even  # noqa: E731
        cs = F(SQRT1_2)
        if kind in ("mjpeg", "mpeg2"):
            M = self.MF
            tmp = [F(0)] * 64
            for v in range(8):
                for x in range(8):
                    s = F(0)
                    for u in range(8):
                        s = r(M[x][u] * c[v * 8 + u] + s)
                    tmp[v * 8 + x] = s
            out = []
            for y in range(8):
                for x in range(8):
                    s = F(0)
                    for v in range(8):
                        s = r(M[y][v] * tmp[v * 8 + x] + s)
                    out.append(r(s + 128) if kind == "mjpeg" else s)
            return out
        if kind == "dv88":
            out = []
            for y in range(8):
                for x in range(8):
                    s = F(0)
                    for v in range(8):
                        for u in range(8):
                            t = r((cs if u == 0 else 1) * (cs if v == 0 else 1))
                            t = r(t * c[v * 8 + u])
                            t = r(t * self.c16F[x * 8 + u])
                            s = r(t * self.c16F[y * 8 + v] + s)
                    out.append(r(s / 4))
            return out
        b = [F(0)] * 64
        for pair in range(4):
            for u in range(8):
                b[pair * 16 + u] = F(c[pair * 16 + u] + c[pair * 16 + 8 + u])
                b[pair * 16 + 8 + u] = F(c[pair * 16 + u] - c[pair * 16 + 8 + u])
        h = [F(0)] * 64
        for row in range(8):
            for x in range(8):
                s = F(0)
                for u in range(8):
                    t = r((cs if u == 0 else 1) * b[row * 8 + u])
                    s = r(t * self.c16F[x * 8 + u] + s)
                h[row * 8 + x] = r(s / 2)
        out = [F(0)] * 64
        for field in range(2):
            for y in range(4):
                for x in range(8):
                    s = F(0)
                    for v in range(4):
                        t = r((cs if v == 0 else 1) * h[(v * 2 + field) * 8 + x])
                        s = r(t * self.c8F[y * 4 + v] + s)
                    out[(y * 2 + field) * 8 + x] = r(s / 2)
        return out

    # ---- exact value and error bound
    def exact(self, kind, c):
        """(exact values as Fractions, bound on |binary64 evaluation - exact| per sample)."""
        cs = F(SQRT1_2)
        if kind in ("mjpeg", "mpeg2"):
            M = self.MF
            tmp = [[sum((M[x][u] * c[v * 8 + u] for u in range(8)), F(0)) for x in range(8)] for v in range(8)]
            mag = [[sum(abs(self.M[x][u] * c[v * 8 + u]) for u in range(8)) for x in range(8)] for v in range(8)]
            vals, bounds = [], []
            for y in range(8):
                for x in range(8):
                    vals.append(sum((M[y][v] * tmp[v][x] for v in range(8)), F(0)) + (128 if kind == "mjpeg" else 0))
                    m = sum(abs(self.M[y][v]) * mag[v][x] for v in range(8))
                    # 8 roundings per row-pass sum (relative to its magnitude), 8 more per column
                    # sum, one for the +128: 2 * 17 u covers them with room for second-order terms
                    bounds.append(34 * U * m + (U * (m + 128) if kind == "mjpeg" else 0) + 1e-300)
            return vals, bounds
        if kind == "dv88":
            vals, bounds = [], []
            for y in range(8):
                for x in range(8):
                    s, m = F(0), 0.0
                    for v in range(8):
                        for u in range(8):
                            t = (cs if u == 0 else 1) * (cs if v == 0 else 1) * c[v * 8 + u] * self.c16F[x * 8 + u] * self.c16F[y * 8 + v]
                            s += t
                            m += abs(float(t))
                    vals.append(s / 4)
                    bounds.append(80 * U * m / 4 + 1e-300)   # 4 products and one sum per term, 64 terms
            return vals, bounds
        b = [F(c[p * 16 + u] + c[p * 16 + 8 + u]) if k == 0 else F(c[p * 16 + u] - c[p * 16 + 8 + u])
             for p in range(4) for k in range(2) for u in range(8)]
        h = [sum(((cs if u == 0 else 1) * b[row * 8 + u] * self.c16F[x * 8 + u] for u in range(8)), F(0)) / 2
             for row in range(8) for x in range(8)]
        hm = [sum(abs(float((cs if u == 0 else 1) * b[row * 8 + u] * self.c16F[x * 8 + u])) for u in range(8)) / 2
              for row in range(8) for x in range(8)]
        vals, bounds = [F(0)] * 64, [0.0] * 64
        for field in range(2):
            for y in range(4):
                for x in range(8):
                    s = sum(((cs if v == 0 else 1) * h[(v * 2 + field) * 8 + x] * self.c8F[y * 4 + v] for v in range(4)), F(0))
                    m = sum(abs(float((cs if v == 0 else 1) * self.c8F[y * 4 + v])) * hm[(v * 2 + field) * 8 + x] for v in range(4))
                    vals[(y * 2 + field) * 8 + x] = s / 2
                    bounds[(y * 2 + field) * 8 + x] = 40 * U * m / 2 + 1e-300
        return vals, bounds
