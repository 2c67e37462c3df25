"""The oracle's DV25 entropy front end and reconstruction on the reference's two test frames
(testdata/dv/valid: one 525/60 4:1:1 frame and one 625/50 4:2:0 frame of FFmpeg's testsrc2).

The reference checks these frames structurally (crates/codecs/dv/tests/reference_vectors.rs:11-60)
and against FFmpeg when that is installed (:33-43), which it is not here, so no decoded golden
frames exist. What pins the front end and the reconstruction:
  * the reference's assertions: both frames parse without DIF issues, every video segment decodes
    with all 30 blocks ending in EOB inside the segment's bits, and the picture is not blank;
    a damaged DIF identifier fails the frame, and with concealment only its segment (0, 0) is
    replaced;
  * the coefficient code is a complete prefix code (Kraft sum exactly 1, 746 entries, signed
    levels, EOB = run 127), so every bit string decodes one way;
  * the macroblock shuffle tiles each picture exactly once;
  * both frames draw the same testsrc2 layout at two heights: the 525/60 luma agrees with the
    625/50 luma resampled to 480 rows (per 16x16 area mean absolute difference 1.8 on average and
    above 10 on 7 % of the areas, the noise-textured checkerboard; a wrong macroblock origin
    table raises these to 32 / 44 %, a wrong sequence offset to 3.8 / 12 %). Shuffle errors
    that move few macroblocks inside uniform areas stay below these bounds; they are what the
    tiling check and the GPU's bit-exact parity with this oracle leave to the reference's
    FFmpeg comparison.
"""
import ctypes as C
import hashlib
import os
from fractions import Fraction

import numpy as np
import pytest

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

HERE = os.path.join(os.path.dirname(__file__), "golden", "dv")
FRAMES = {"dv25-525-60-one-frame.dv": abi.DV_SYSTEM_525_60, "dv25-625-50-one-frame.dv": abi.DV_SYSTEM_625_50}


def _binding():
    from oracle import binding

    return binding


def read(name):
    with open(os.path.join(HERE, name), "rb") as f:
        return f.read()


def decode(oracle, f):
    u = wl.dv_unit(f["system"], f["levels"], f["info"])
    out = u.host_outputs()
    s = u.host_struct(out)
    assert oracle.oracle_dv_reconstruct(C.byref(s)) == 0
    return out


def test_frame_fixtures_match_reference_checksums():
    with open(os.path.join(HERE, "SHA256SUMS")) as f:
        want = dict(reversed(line.split()) for line in f)
    want = {os.path.basename(k): v for k, v in want.items()}
    assert set(want) == set(FRAMES)
    for name, digest in want.items():
        assert hashlib.sha256(read(name)).hexdigest() == digest, name


@pytest.mark.parametrize("name", sorted(FRAMES))
def test_frames_decode_completely(oracle, name):
    f = _binding().dv_frame_levels(read(name))
    assert f["system"] == FRAMES[name]
    seqs = 12 if f["system"] == abi.DV_SYSTEM_625_50 else 10
    assert f["levels"].shape == (seqs * 810, 64)
    assert f["concealed"] == []
    assert (f["levels"][:, 1:] != 0).any()
    assert (np.abs(f["levels"][:, 0]) <= 256).all()
    out = decode(oracle, f)
    assert out["y"].shape == ((576, 720) if seqs == 12 else (480, 720))
    assert (out["y"] != 0).any() and out["y"].std() > 20


def test_vlc_table_is_a_complete_prefix_code():
    t = _binding().dv_vlc_table()
    assert len(t["code"]) == 746
    assert sum(Fraction(1, 2 ** int(n)) for n in t["len"]) == 1
    codes = sorted(format(int(c), "0%db" % int(n)) for c, n in zip(t["code"], t["len"]))
    assert all(not b.startswith(a) for a, b in zip(codes, codes[1:]))
    assert int(t["len"].max()) == 16
    # decoder.rs:620-625: signed entries and an EOB (run 127)
    assert (t["level"] < 0).any() and ((t["run"] == 127) & (t["level"] == 0)).sum() == 1
    # a nonzero level has two codes, + then -, that differ in the last (sign) bit only
    for i in np.nonzero(t["level"] < 0)[0]:
        assert (t["run"][i - 1], t["level"][i - 1], t["len"][i - 1]) == (t["run"][i], -t["level"][i], t["len"][i])
        assert int(t["code"][i - 1]) | 1 == int(t["code"][i]) and int(t["code"][i - 1]) & 1 == 0


@pytest.mark.parametrize("system", [abi.DV_SYSTEM_525_60, abi.DV_SYSTEM_625_50])
def test_macroblock_shuffle_tiles_the_picture(oracle, system):
    """decoder.rs:643-657 asks for distinct origins; here every luma sample is covered once:
    625/50 macroblocks are 16x16, 525/60 ones 32x8 except 16x16 at the right edge (x >= 704)."""
    seqs, H = (12, 576) if system == abi.DV_SYSTEM_625_50 else (10, 480)
    cover = np.zeros((H, 720), np.int32)
    xy = (C.c_int * 10)()
    for seq in range(seqs):
        for slot in range(27):
            oracle.oracle_dv_macroblock_coordinates(system, seq, slot, xy)
            for m in range(5):
                x, y = xy[2 * m] * 8, xy[2 * m + 1] * 8
                w, h = (16, 16) if system == abi.DV_SYSTEM_625_50 or x >= 704 else (32, 8)
                cover[y:y + h, x:x + w] += 1
    assert (cover == 1).all()


def test_damaged_identifier_fails_and_conceals_one_segment(oracle):
    """reference_vectors.rs:45-60: byte 7 * 80 (first video block of sequence 0) rewritten."""
    b = _binding()
    clean = read("dv25-525-60-one-frame.dv")
    damaged = bytearray(clean)
    damaged[7 * 80] = 0x76
    with pytest.raises(b.StreamError):
        b.dv_frame_levels(bytes(damaged))
    f = b.dv_frame_levels(bytes(damaged), conceal=True)
    assert f["concealed"] == [(0, 0)]
    got, want = decode(oracle, f), decode(oracle, b.dv_frame_levels(clean))
    xy = (C.c_int * 10)()
    oracle.oracle_dv_macroblock_coordinates(abi.DV_SYSTEM_525_60, 0, 0, xy)
    mask_y = np.zeros(got["y"].shape, bool)
    mask_c = np.zeros(got["cb"].shape, bool)
    for m in range(5):
        x, y = xy[2 * m] * 8, xy[2 * m + 1] * 8
        if x >= 704:
            mask_y[y:y + 16, x:x + 16] = True
            mask_c[y:y + 16, x // 4:x // 4 + 4] = True
        else:
            mask_y[y:y + 8, x:x + 32] = True
            mask_c[y:y + 8, x // 4:x // 4 + 8] = True
    assert (got["y"][mask_y] == 16).all() and np.array_equal(got["y"][~mask_y], want["y"][~mask_y])
    for n in ("cb", "cr"):
        assert (got[n][mask_c] == 128).all() and np.array_equal(got[n][~mask_c], want[n][~mask_c])


def test_front_end_rejects_bad_frames():
    b = _binding()
    data = read("dv25-625-50-one-frame.dv")
    for bad in (data[:3], data[:-80], data + bytes(80), bytes([0x80]) + data[1:]):
        with pytest.raises(b.StreamError):
            b.dv_frame_levels(bad)


def test_525_and_625_frames_agree(oracle):
    b = _binding()
    ntsc = decode(oracle, b.dv_frame_levels(read("dv25-525-60-one-frame.dv")))["y"].astype(int)
    pal = decode(oracle, b.dv_frame_levels(read("dv25-625-50-one-frame.dv")))["y"].astype(int)
    rows = np.minimum(((np.arange(480) + 0.5) * 576 / 480).astype(int), 575)
    d = np.abs(ntsc - pal[rows]).reshape(30, 16, 45, 16).mean(axis=(1, 3))
    d[:3, :7] = 0   # the timecode box keeps its size at both heights
    assert d.mean() <= 2.5, d.mean()
    assert (d > 10).mean() <= 0.09, (d > 10).mean()
