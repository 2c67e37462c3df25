"""The oracle's H.264 CAVLC front end (oracle/h264_entropy.c), which brings H.264 bitstreams to the
device path, pinned by:
* the CAVLC code tables: every table a prefix code (Kraft sum at most 1), and every entry of every
  coeff_token / total_zeros / run_before table round-tripping between this suite's own encoder
  (tests/h264_bitstream.py) and the oracle's decoder, including level escapes;
* the reference's CAVLC unit vectors (crates/codecs/h264/src/cavlc.rs:516-578);
* the reference's weighted bi-prediction known answers through whole bitstreams
  (crates/codecs/h264/src/decoder.rs:16853-16950: explicit [88, 105, 150], implicit [57, 95, 150]);
* seeded multi-macroblock, multi-slice I / P / B streams of every supported macroblock type
  (I_PCM, Intra4x4, Intra8x8, Intra16x16, P and B 16x16 / 16x8 / 8x16 / 8x8 with every sub type,
  P_Skip), whose decoded records must equal the coded syntax: level records, intra modes, coded
  block patterns, QPs, slices, reference indices, mvds, and motion vectors against an independent
  restatement of the motion vector predictor (8.4.1.3, P_Skip 8.4.1.1);
* planted front-end mistakes (nC rounding, the neighbour rule of the predictor, the intra mode
  predictor, the level suffix growth), each failing one of these."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from mmrecode_amd import abi
from oracle import binding as ob

import h264_bitstream as hb
import h264_streams as hs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def decode_block(lib, bits_bytes, nc, n):
    buf = (C.c_uint8 * max(len(bits_bytes), 1)).from_buffer_copy(bits_bytes or b"\0")
    out = (C.c_int32 * n)()
    used = C.c_size_t()
    t = lib.oracle_h264_cavlc_block(buf, len(bits_bytes), nc, n, out, C.byref(used))
    return t, list(out), used.value


def encode(nc, coeffs):
    w = hb.BitWriter()
    total = hb.encode_block(w, nc, coeffs)
    nbits = len(w.bits)
    bits = w.bits + [0] * (-len(w.bits) % 8)
    data = bytes(int("".join(map(str, bits[i:i + 8])), 2) for i in range(0, len(bits), 8))
    return total, data, nbits


# ------------------------------------------------------------------ tables and blocks
def test_cavlc_code_tables_are_prefix_codes():
    tables = [hb.COEFF_TOKEN[k] for k in range(4)] + [hb.CHROMA_DC_TOKEN] + hb.TOTAL_ZEROS + hb.CHROMA_DC_TOTAL_ZEROS + hb.RUN_BEFORE
    for lengths, codes in tables:
        ok, kraft = hb.prefix_code_ok(lengths, codes)
        assert ok and kraft <= 1.0 + 1e-12, (lengths, codes)


def test_cavlc_reference_vectors():
    """cavlc.rs:516-546: an empty block, one trailing one, a non-trailing +2, a run of three zeros."""
    L = ob.lib()
    t, c, _ = decode_block(L, bytes([0b1000_0000]), 0, 16)
    assert t == 0 and c == [0] * 16
    t, c, _ = decode_block(L, bytes([0b0101_0000]), 0, 16)
    assert t == 1 and c[0] == 1 and not any(c[1:])
    t, c, _ = decode_block(L, bytes([0b0001_0111]), 0, 16)
    assert c[0] == 2 and sum(map(abs, c)) == 2
    t, c, _ = decode_block(L, bytes([0b0100_0110]), 0, 16)
    assert c[3] == 1 and sum(c) == 1
    for nc, coeffs in [(0, [0, 2, 0, -1, 1, 0, 0, -7, 0, 0, 0, 0, 0, 0, 0, 3]), (2, [12, -20, 0, 0, 0, 1, -1, 1] + [0] * 8),
                       (4, [0] * 15 + [-2]), (8, [1] * 16), (3, [0, -3, 0, 1, 0, 0, 8, 0, -1, 0, 0, 2, 0, 0, 1]),
                       (-1, [3, 0, -2, 1])]:
        total, data, nbits = encode(nc, coeffs)
        t, c, used = decode_block(L, data, nc, len(coeffs))
        assert c == coeffs and t == total == sum(1 for v in coeffs if v) and used == nbits


@pytest.mark.parametrize("nc", [-1, 0, 1, 2, 3, 4, 7, 8, 16])
def test_cavlc_round_trip_every_table_entry(nc):
    """Every (TotalCoeff, TrailingOnes) of the nC class, every total_zeros and run_before, and
    levels up to the escape codes."""
    L = ob.lib()
    rng = np.random.default_rng(100 + nc)
    sizes = [4] if nc == -1 else [15, 16]
    for n in sizes:
        for total in range(0, n + 1):
            for t1 in range(0, min(3, total) + 1):
                for trial in range(6):
                    pos = sorted(rng.choice(n, total, replace=False).tolist())
                    coeffs = [0] * n
                    for j, p in enumerate(reversed(pos)):   # highest frequency first
                        if j < t1:
                            coeffs[p] = int(rng.choice([-1, 1]))
                        else:
                            mag = int(rng.integers(2 if j == t1 else 1, 6))
                            if trial >= 4:
                                mag = int(rng.integers(1, 5000))
                            if j == t1 and t1 < 3 and mag == 1:
                                mag = 2
                            coeffs[p] = mag if rng.random() < 0.5 else -mag
                    if t1 < 3 and total > t1 and abs(coeffs[pos[-1 - t1]]) == 1:
                        coeffs[pos[-1 - t1]] *= 2
                    total_w, data, nbits = encode(nc, coeffs)
                    t, c, used = decode_block(L, data, nc, n)
                    assert c == coeffs and t == total_w and used == nbits, (nc, n, total, t1, coeffs)


# ------------------------------------------------------------------ bitstreams
@pytest.mark.parametrize("mode, want", [("explicit", (88, 105, 150)), ("implicit", (57, 95, 150))])
def test_reference_weighted_bipred_known_answers(mode, want):
    L = ob.lib()
    decs = hs.decode_stream(L, hs.weighted_bipred_stream(mode))
    assert [d.info.slice_type for d in decs] == [2, 0, 1]
    units = [hs.unit_of(d) for d in decs]
    planes = hs.reconstruct_on_oracle(L, units)
    b = planes[units[-1].meta["slot"]]
    for n, v in zip(("y", "cb", "cr"), want):
        assert (b[n] == v).all(), (mode, n, np.unique(b[n]))


def check_stream(lib, aus, expected):
    """Decoded records against the coded syntax. Returns a list of mismatch descriptions."""
    fe = hs.FrontEnd(lib)
    bad = []
    try:
        pics = []
        for pi, au in enumerate(aus):
            d = fe.decode(au)
            if isinstance(d, int) or d is None:
                return [f"picture {pi}: front end status {d}"]
            pics.append(d)
    finally:
        fe.close()
    for pi, (d, exp) in enumerate(zip(pics, expected)):
        mbs, blocks, levels = d.inputs["mbs"], d.inputs["blocks"], d.inputs["levels"]
        intra = d.inputs.get("intra")
        for mb, e in enumerate(exp):
            m = mbs[mb]
            where = f"picture {pi} mb {mb} ({e.get('kind', 'skip' if e.get('skip') else 'inter')})"
            if bool(m["flags"] & abi.H264_MB_INTRA) != e["intra"] or int(m["cbp"]) != e["cbp"] or int(m["qp"]) != e["qp"] or \
                    int(m["slice_id"]) != e["slice"]:
                bad.append(f"{where}: flags/cbp/qp/slice {m} vs {e['intra'], e['cbp'], e['qp'], e['slice']}")
                continue
            if bool(m["flags"] & abi.H264_MB_TRANSFORM_8X8) != bool(e.get("t8", False)):
                bad.append(f"{where}: transform_size_8x8")
            rec = e.get("record")
            if rec is not None and not np.array_equal(levels[m["coef_offset"]], rec):
                bad.append(f"{where}: level record")
            if e["intra"]:
                r = intra[mb]
                if int(r["type"]) != e["type"]:
                    bad.append(f"{where}: intra type")
                elif e["type"] == abi.H264_INTRA_16X16 and (int(r["luma_mode"]) != e["luma_mode"] or int(r["chroma_mode"]) != e["chroma_mode"]):
                    bad.append(f"{where}: Intra16x16 modes")
                elif e["type"] in (abi.H264_INTRA_4X4, abi.H264_INTRA_8X8):
                    k = 4 if e["type"] == abi.H264_INTRA_8X8 else 16
                    if list(r["modes"][:k]) != e["modes"] or int(r["chroma_mode"]) != e["chroma_mode"]:
                        bad.append(f"{where}: intra modes {list(r['modes'][:k])} vs {e['modes']}")
                continue
            if not np.array_equal(blocks[mb]["ref_idx"].T, e["ref"]):
                bad.append(f"{where}: ref_idx")
            if not np.array_equal(d.mvd[mb], e["mvd"]):
                bad.append(f"{where}: mvd")
            mv = np.where(e["ref"][..., None] >= 0, e["mv"], 0)
            if not np.array_equal(np.transpose(blocks[mb]["mv"], (1, 0, 2)), mv):
                bad.append(f"{where}: motion vectors {blocks[mb]['mv'].tolist()} vs {mv.tolist()}")
    return bad


STREAMS = [(1, 5, 4, "IPBPB", True), (2, 3, 3, "IPPBB", False), (3, 7, 2, "IPBPBB", True), (4, 1, 1, "IPB", True),
           (5, 4, 5, "IBPB", True)]


@pytest.mark.parametrize("seed, w, h, pattern, high", STREAMS)
def test_random_streams_decode_to_the_coded_syntax(seed, w, h, pattern, high):
    L = ob.lib()
    aus, expected = hs.random_stream(seed, w, h, pattern if "B" not in pattern[:2] else pattern, high=high)
    bad = check_stream(L, aus, expected)
    assert not bad, bad[:5]
    units = [hs.unit_of(d) for d in hs.decode_stream(L, aus)]
    hs.reconstruct_on_oracle(L, units)


# ------------------------------------------------------------------ planted mistakes
PLANTS = {
    "nc_rounding": ("if (A.avail && B.avail) return (na + nbv + 1) >> 1;", "if (A.avail && B.avail) return (na + nbv) >> 1;"),
    "no_copy_of_a": ("    if (!B.avail && !C.avail && A.avail) {\n        rb = rc = ra;",
                     "    if (0) {\n        rb = rc = ra;"),
    "mode_pred_max": ("return ma < mbv ? ma : mbv;", "return ma > mbv ? ma : mbv;"),
    "suffix_growth": ("if ((lv < 0 ? -lv : lv) > (3 << (suffix_length - 1)) && suffix_length < 6) ++suffix_length;",
                      "if ((lv < 0 ? -lv : lv) > (2 << (suffix_length - 1)) && suffix_length < 6) ++suffix_length;"),
}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    base = tmp_path_factory.mktemp("h264_planted")
    libs = {}
    for name, (old, new) in PLANTS.items():
        d = base / name
        (d / "oracle").mkdir(parents=True)
        (d / "include").mkdir()
        shutil.copy(os.path.join(ROOT, "include", "mmr_recon.h"), d / "include")
        srcs = []
        for f in os.listdir(os.path.join(ROOT, "oracle")):
            if f.endswith((".c", ".h")):
                text = open(os.path.join(ROOT, "oracle", f)).read()
                if f == "h264_entropy.c":
                    assert old in text, name
                    text = text.replace(old, new)
                (d / "oracle" / f).write_text(text)
                if f.endswith(".c"):
                    srcs.append(str(d / "oracle" / f))
        so = d / "liboracle.so"
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c11", "-D_GNU_SOURCE", "-shared", "-o",
                        str(so), *srcs, "-lm"], check=True)
        L = C.CDLL(str(so))
        L.oracle_h264_decoder_new.restype = C.c_void_p
        L.oracle_h264_decoder_free.argtypes = [C.c_void_p]
        L.oracle_h264_decode_access_unit.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        L.oracle_h264_picture_records.argtypes = [C.c_void_p] * 8
        libs[name] = L
    return libs


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_front_end_mistakes_fail(planted, plant):
    L = planted[plant]
    for seed, w, h, pattern, high in STREAMS:
        aus, expected = hs.random_stream(seed, w, h, pattern, high=high)
        if check_stream(L, aus, expected):
            return
    pytest.fail(f"planted mistake {plant} went unnoticed")
