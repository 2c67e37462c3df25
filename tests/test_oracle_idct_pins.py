"""Exactness pins for the oracle's f64 IDCTs (MJPEG / MPEG-2 inverse_dct, DV inverse_88 / inverse_248)
and MPEG-2 dequantisation, chosen so that plausible restatement mistakes fail:

* the basis table is the cosine basis to its printed digits (and, where the reference is mounted,
  its literal table value for value); the DV cosine tables are the C library's cos of the reference's
  argument expression;
* exact rationals: on seeded blocks every sample whose exact value (the real sum over the same
  binary64 constants) lies farther from a half-integer than the proven binary64 error bound must
  come out as round-half-away of that exact value, shifted and clamped; this fixes the rounding
  direction, the level shift, the clamps, the basis orientation and every term, whatever the
  evaluation order (samples inside the window are counted);
* IEEE-754 order: every sample, including those in the window, equals the reference's loop order in
  binary64 (tests/idct_reference.py), on the seeded blocks and on tests/golden/idct_pins.json's
  blocks, where an exact half-way result (tie) or a fused multiply-add changes a sample;
* planted mistakes: the oracle rebuilt with floor, trunc or round-half-to-even in place of round,
  or with its products fused into the sums (fma), fails the checks above;
* H.262 7.4.2.3 / 7.4.4 for intra and non-intra dequantisation with intra_dc_precision 0-3, with a
  planted /32 in the intra branch failing it.
"""
import ctypes as C
import json
import math
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

from oracle import binding as ob

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)

from idct_reference import Transforms, basis, round_half_away  # noqa: E402

PINS = json.load(open(os.path.join(HERE, "golden", "idct_pins.json")))
REFERENCE = "/root/reference"


def _unpack(d):
    c = [0] * 64
    for k, v in d.items():
        c[int(k)] = int(v)
    return c


def _blocks(kind, n, seed):
    """Seeded coefficient blocks in each transform's input range: sparse Laplacian, dense, extremes."""
    rng = np.random.default_rng([seed, Transforms.KINDS.index(kind)])
    lim = 2047 if kind in ("mjpeg", "mpeg2") else 1500
    out = []
    for i in range(n):
        c = np.zeros(64, np.int64)
        if i % 5 == 4:    # dense
            c[:] = rng.integers(-200, 201, 64)
        else:
            idx = rng.choice(64, int(rng.integers(1, 14)), replace=False)
            c[idx] = np.round(rng.laplace(0, rng.choice([4, 30, 300]), idx.size))
        if i % 7 == 6:    # extremes
            c[rng.choice(64, 3, replace=False)] = rng.choice([-lim, lim], 3)
        if kind.startswith("dv"):
            c[0] = rng.integers(0, 2048)
        elif i % 3 == 0:
            c[0] = rng.integers(-1024, 1024)
        out.append([int(v) for v in np.clip(c, -lim, lim)])
    return out


def _fixture_blocks(kind):
    return [_unpack(d) for d in PINS[kind]["tie"] + PINS[kind]["fused"]]


COUNTS = {"mjpeg": 90, "mpeg2": 90, "dv88": 24, "dv248": 60}


@pytest.fixture(scope="module")
def T():
    return Transforms(ob.lib())


@pytest.fixture(scope="module")
def expected(T):
    """Per kind: blocks, exact values + bounds (random blocks) and the reference-order samples."""
    out = {}
    for kind in Transforms.KINDS:
        rand = _blocks(kind, COUNTS[kind], 5)
        exact = [T.exact(kind, c) for c in rand]
        blocks = rand + _fixture_blocks(kind)
        ieee = [[T.finish(kind, v) for v in T.plain(kind, c)] for c in blocks]
        out[kind] = (rand, exact, blocks, ieee)
    return out


def _run(lib, kind, c):
    a = np.ascontiguousarray(c, dtype=np.int32)
    if kind == "mjpeg":
        o = np.zeros(64, np.uint8)
        lib.oracle_mjpeg_inverse_dct(a.ctypes.data_as(C.c_void_p), o.ctypes.data_as(C.c_void_p))
    elif kind == "mpeg2":
        o = np.zeros(64, np.int16)
        lib.oracle_mpeg2_inverse_dct(a.ctypes.data_as(C.c_void_p), o.ctypes.data_as(C.c_void_p))
    else:
        o = np.zeros(64, np.uint8)
        fn = lib.oracle_dv_inverse_88 if kind == "dv88" else lib.oracle_dv_inverse_248
        fn(a.ctypes.data_as(C.c_void_p), o.ctypes.data_as(C.c_void_p))
    return [int(v) for v in o]


def _exact_mismatches(lib, kind, exp):
    rand, exact, _, _ = exp
    bad, window = 0, 0
    for c, (vals, bounds) in zip(rand, exact):
        got = _run(lib, kind, c)
        for i, (q, b) in enumerate(zip(vals, bounds)):
            dist = abs(abs(q - math.floor(q)) - 0.5)
            if dist <= b:
                window += 1
                continue
            if got[i] != Transforms.finish(kind, q):
                bad += 1
    return bad, window


def _ieee_mismatches(lib, kind, exp):
    _, _, blocks, ieee = exp
    return sum(g != w for c, want in zip(blocks, ieee) for g, w in zip(_run(lib, kind, c), want))


# ------------------------------------------------------------------ tables
def test_basis_is_the_cosine_basis_to_its_printed_digits():
    """The literals carry 16 significant digits, so they sit within a few 1e-15 of the cosines; a
    transposed or mis-signed entry is off by far more."""
    M = basis()
    for x in range(8):
        for u in range(8):
            want = (math.sqrt(0.5) if u == 0 else 1.0) * 0.5 * math.cos((2 * x + 1) * u * math.pi / 16)
            assert abs(M[x][u] - want) <= 5e-15, (x, u, M[x][u], want)


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="the reference source is not mounted here")
def test_basis_equals_the_reference_literal_tables():
    """IDCT_MATRIX (crates/codecs/mjpeg/src/transform.rs:2-84) and MATRIX
    (crates/codecs/mpeg2/src/transform.rs:6-87), value for value."""
    M = [v for row in basis() for v in row]
    for path, name in (("crates/codecs/mjpeg/src/transform.rs", "IDCT_MATRIX"), ("crates/codecs/mpeg2/src/transform.rs", "MATRIX")):
        src = open(os.path.join(REFERENCE, path)).read()
        body = src[src.index(f"const {name}:"):]
        body = body[body.index("= [") + 2:body.index("\n];")]
        vals = [float(v.replace("_", "")) for v in re.findall(r"-?\d[\d_]*\.[\d_]+(?:[eE]-?\d+)?", body)]
        assert vals == M, path


def test_dv_cosine_tables_are_c_library_cos_of_the_reference_expression(T):
    """cos((2x + 1) as f64 * u as f64 * PI / 16.0) and the 8-point table with / 8.0
    (crates/codecs/dv/src/decoder.rs:520-602); Python's math.cos is the same C library cos."""
    for x in range(8):
        for u in range(8):
            assert T.c16[x * 8 + u] == math.cos(float(2 * x + 1) * float(u) * math.pi / 16.0), (x, u)
    for y in range(8):
        for v in range(4):
            assert T.c8[y * 4 + v] == math.cos(float(2 * y + 1) * float(v) * math.pi / 8.0), (y, v)


# ------------------------------------------------------------------ the oracle
@pytest.mark.parametrize("kind", Transforms.KINDS)
def test_exact_rational_outside_the_error_window(expected, kind):
    bad, window = _exact_mismatches(ob.lib(), kind, expected[kind])
    total = sum(len(v[0]) for v in expected[kind][1])
    assert bad == 0
    assert window <= total // 200, f"{window} of {total} samples within the error bound of a half-integer"


@pytest.mark.parametrize("kind", Transforms.KINDS)
def test_reference_order_in_binary64_on_every_sample(expected, kind):
    assert _ieee_mismatches(ob.lib(), kind, expected[kind]) == 0


@pytest.mark.parametrize("kind", Transforms.KINDS)
def test_fixture_blocks_decide_ties_and_fusion(T, kind):
    """The committed blocks have teeth: each tie block holds an exact half-integer in the reference's
    binary64 evaluation; each fused block has a sample that FMA contraction rounds differently."""
    assert PINS[kind]["tie"], kind
    for d in PINS[kind]["tie"]:
        vals = T.plain(kind, _unpack(d))
        assert any(v - math.floor(v) == 0.5 for v in vals)
    for d in PINS[kind]["fused"]:
        c = _unpack(d)
        plain, fused = T.plain(kind, c), T.fused(kind, c)
        assert any(round_half_away(a) != round_half_away(b) for a, b in zip(plain, fused))


# ------------------------------------------------------------------ planted mistakes
ORACLE_SOURCES = ["mjpeg.c", "mpeg2.c", "dv.c", "h264.c", "h264_intra.c", "deblock.c", "compose.c", "jpeg_entropy.c",
                  "mpeg2_entropy.c", "dv_entropy.c"]


def _fuse(src):
    """Products fused into the accumulations that follow them."""
    n0 = src.count("fma(")
    src = re.sub(r"double product = ([^;]+?) \* ([^;]+?);\s*(\w+(?:\[[^\]]+\])?) = \3 \+ product;", r"\3 = fma(\1, \2, \3);", src)
    src = re.sub(r"term = term \* ([^;]+?);\s*sum = sum \+ term;", r"sum = fma(term, \1, sum);", src)
    src = re.sub(r"term = term \* (COS8\[[^;]+?\]);\s*sum = sum \+ term;", r"sum = fma(term, \1, sum);", src)
    assert src.count("fma(") > n0
    return src


PLANTS = {
    "floor": lambda s: s.replace("round(", "floor("),
    "trunc": lambda s: s.replace("round(", "trunc("),
    "half_even": lambda s: s.replace("round(", "rint("),
    "fused": _fuse,
}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    base = tmp_path_factory.mktemp("planted")
    libs = {}
    for name, edit in PLANTS.items():
        d = base / name
        (d / "oracle").mkdir(parents=True)
        (d / "include").mkdir()
        shutil.copy(os.path.join(ROOT, "include", "mmr_recon.h"), d / "include")
        for f in os.listdir(os.path.join(ROOT, "oracle")):
            if f.endswith((".c", ".h")):
                text = open(os.path.join(ROOT, "oracle", f)).read()
                if f in ("mjpeg.c", "mpeg2.c", "dv.c"):
                    text = edit(text)
                (d / "oracle" / f).write_text(text)
        so = d / "liboracle_planted.so"
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c11", "-D_GNU_SOURCE", "-shared",
                        "-o", str(so), *[str(d / "oracle" / f) for f in ORACLE_SOURCES], "-lm"], check=True)
        libs[name] = C.CDLL(str(so))
    return libs


@pytest.mark.parametrize("plant", ["floor", "trunc"])
@pytest.mark.parametrize("kind", Transforms.KINDS)
def test_planted_rounding_direction_fails_exact_check(planted, expected, plant, kind):
    bad, _ = _exact_mismatches(planted[plant], kind, expected[kind])
    assert bad > 0


@pytest.mark.parametrize("kind", Transforms.KINDS)
def test_planted_half_to_even_fails_on_ties(planted, expected, kind):
    assert _ieee_mismatches(planted["half_even"], kind, expected[kind]) > 0


@pytest.mark.parametrize("kind", [k for k in Transforms.KINDS if PINS[k]["fused"]])
def test_planted_fused_multiply_add_fails(planted, expected, kind):
    assert _ieee_mismatches(planted["fused"], kind, expected[kind]) > 0


# ------------------------------------------------------------------ MPEG-2 dequantisation
ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
          13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59,
          52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63]
ALTERNATE = [0, 8, 16, 24, 1, 9, 2, 10, 17, 25, 32, 40, 48, 56, 57, 49, 41, 33, 26, 18, 3, 11, 4, 12, 19,
             27, 34, 42, 50, 58, 35, 43, 51, 59, 20, 28, 5, 13, 6, 14, 21, 29, 36, 44, 52, 60, 37, 45, 53,
             61, 22, 30, 7, 15, 23, 31, 38, 46, 54, 62, 39, 47, 55, 63]


def _div_trunc(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def h262_dequant(levels, W, qscale, intra, dc_precision, alternate):
    """ITU-T H.262 7.4.2.3 (reconstruction formulae, "/" = integer division truncating toward zero),
    7.4.3 (saturation to [-2048, 2047]) and 7.4.4 (mismatch control on F'[7][7]); levels by scan
    position, natural order out (7.3 inverse scan)."""
    scan = ALTERNATE if alternate else ZIGZAG
    F = [0] * 64
    for p in range(64):
        qf = int(levels[p])
        pos = scan[p]
        if intra and p == 0:
            v = (8 >> dc_precision) * qf
        elif intra:
            v = _div_trunc(qf * int(W[pos]) * qscale * 2, 32)
        else:
            k = 0 if qf == 0 else (1 if qf > 0 else -1)
            v = _div_trunc((2 * qf + k) * int(W[pos]) * qscale, 32)
        F[pos] = max(-2048, min(2047, v))
    if sum(F) % 2 == 0:
        F[63] += -1 if F[63] % 2 else 1
    return F


def _dequant_checks(lib):
    rng = np.random.default_rng(9)
    bad = 0
    for trial in range(400):
        lv = np.zeros(64, np.int16)
        idx = rng.choice(63, int(rng.integers(1, 20)), replace=False) + 1
        lv[idx] = np.round(rng.laplace(0, rng.choice([3, 40, 400]), idx.size)).clip(-2048, 2047)
        W = rng.integers(1, 256, 64).astype(np.uint8)
        qs = int(rng.integers(1, 113))
        alt = int(rng.integers(0, 2))
        out = np.zeros(64, np.int32)
        if trial % 2 == 0:
            prec = trial // 2 % 4
            lv[0] = rng.integers(0, 1 << (8 + prec))
            lib.oracle_mpeg2_dequantize_intra(lv.ctypes.data_as(C.c_void_p), 0, qs, prec, alt, W.ctypes.data_as(C.c_void_p),
                                              out.ctypes.data_as(C.c_void_p))
            want = h262_dequant(lv, W, qs, True, prec, alt)
        else:
            lib.oracle_mpeg2_dequantize_non_intra(lv.ctypes.data_as(C.c_void_p), qs, alt, W.ctypes.data_as(C.c_void_p),
                                                  out.ctypes.data_as(C.c_void_p))
            want = h262_dequant(lv, W, qs, False, 0, alt)
        bad += list(out) != want
    return bad


def test_mpeg2_dequantisation_is_h262_with_every_dc_precision():
    assert _dequant_checks(ob.lib()) == 0


def test_planted_intra_divisor_fails_h262(tmp_path):
    d = tmp_path
    (d / "oracle").mkdir()
    (d / "include").mkdir()
    shutil.copy(os.path.join(ROOT, "include", "mmr_recon.h"), d / "include")
    for f in os.listdir(os.path.join(ROOT, "oracle")):
        if f.endswith((".c", ".h")):
            text = open(os.path.join(ROOT, "oracle", f)).read()
            if f == "mpeg2.c":
                n = text.count("(int32_t)matrix[position] / 16;")
                assert n == 1
                text = text.replace("(int32_t)matrix[position] / 16;", "(int32_t)matrix[position] / 32;")
            (d / "oracle" / f).write_text(text)
    so = d / "liboracle_planted.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-std=c11", "-D_GNU_SOURCE", "-shared", "-o", str(so),
                    *[str(d / "oracle" / f) for f in ORACLE_SOURCES], "-lm"], check=True)
    assert _dequant_checks(C.CDLL(str(so))) > 0
