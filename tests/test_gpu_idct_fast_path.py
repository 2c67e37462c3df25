"""The MPEG-2 kernel's fast IDCT falls back to the reference's order where it must: intra blocks
that dequantise exactly to the tie / FMA fixture blocks of tests/golden/idct_pins.json ("mpeg2_odd":
coefficient sums odd, so mismatch control leaves them unchanged; quantiser matrix 16, quantiser
scale 1, intra_dc_precision 3 make every coefficient equal to its level), mixed with random blocks,
reconstruct bit for bit like the oracle. On these blocks the fused evaluation rounds some samples
differently (checked here on the CPU), so a missing or wrong fallback would show."""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
PINS = json.load(open(os.path.join(HERE, "golden", "idct_pins.json")))
ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
          13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59,
          52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63]


def _fixture_coefficients():
    out = []
    for cat in ("tie", "fused"):
        for d in PINS["mpeg2_odd"][cat]:
            c = [0] * 64
            for k, v in d.items():
                c[int(k)] = v
            assert sum(c) % 2 == 1
            out.append(c)
    return out


def test_fixtures_need_the_fallback(oracle):
    from idct_reference import Transforms, round_half_away

    T = Transforms(oracle)
    for c in (c for cat in ("fused",) for c in _fixture_coefficients()[len(PINS["mpeg2_odd"]["tie"]):]):
        plain, fused = T.plain("mpeg2", c), T.fused("mpeg2", c)
        assert any(round_half_away(a) != round_half_away(b) for a, b in zip(plain, fused))


def test_fast_idct_fallback_bit_exact(ctx, oracle):
    fx = _fixture_coefficients()
    rng = np.random.default_rng(17)
    mb_w, mb_h = 6, 2                      # 12 macroblocks, 72 blocks: fixtures first, random after
    n_mb = mb_w * mb_h
    blocks = np.zeros((n_mb * 6, 64), np.int16)
    for b in range(n_mb * 6):
        if b < len(fx):
            coef = fx[b]
        else:
            coef = [0] * 64
            for i in rng.choice(64, int(rng.integers(1, 8)), replace=False):
                coef[int(i)] = int(rng.integers(-300, 301))
            coef[0] = int(rng.integers(0, 2048))
        blocks[b] = [coef[ZIGZAG[p]] for p in range(64)]   # levels by scan position
    mbs = np.zeros(n_mb, dtype=wl.MPEG2_MB_DTYPE)
    mbs["flags"] = abi.MPEG2_MB_INTRA
    mbs["cbp"] = 0x3F
    mbs["qscale"] = 1
    mbs["block_offset"] = np.arange(n_mb) * 6
    matrices = np.full((4, 64), 16, np.uint8)
    u = wl.mpeg2_unit(mb_w * 16, mb_h * 16, abi.MPEG2_PICTURE_I, matrices, mbs, blocks, intra_dc_precision=3)
    d = wl.DeviceUnit(ctx, u)
    ctx.wait(ctx.submit_mpeg2([d.struct]))
    out = u.host_outputs()
    assert oracle.oracle_mpeg2_reconstruct(C.byref(u.host_struct(out))) == 0
    for n in ("y", "cb", "cr"):
        assert np.array_equal(d.download(n), out[n]), n
    d.free()


def test_mjpeg_fast_idct_fallback_bit_exact(ctx, oracle):
    """MJPEG: the "mjpeg" tie / FMA fixture blocks (quantiser 1, so each coefficient is its level),
    blocks with a row past the fast form's coefficient bound (level x quantiser in the millions,
    which must take the reference order) and random blocks, in one grayscale frame, against the
    oracle bit for bit."""
    from idct_reference import Transforms, round_half_away

    T = Transforms(oracle)
    fx = []
    for cat in ("tie", "fused"):
        for d in PINS["mjpeg"][cat]:
            c = [0] * 64
            for k, v in d.items():
                c[int(k)] = v
            fx.append(c)
    assert any(any(round_half_away(a) != round_half_away(b) for a, b in zip(T.plain("mjpeg", c), T.fused("mjpeg", c)))
               for c in fx[len(PINS["mjpeg"]["tie"]):])
    rng = np.random.default_rng(23)
    u = wl.mjpeg_frame(5, 0, width=64, height=24, sampling=((1, 1),), rgb=False)
    levels = u.inputs["levels"]
    quant = u.inputs["quant"]
    quant[0] = 1
    n = levels.shape[0]
    assert n == 24 and len(fx) + 4 <= n
    for b in range(n):
        if b < len(fx):
            coef = fx[b]
        elif b < len(fx) + 4:                   # past the row bound: DC 2047 and two large ACs in row 0
            coef = [0] * 64
            coef[0] = 2047
            for i in rng.choice(np.arange(1, 8), 2, replace=False):
                coef[int(i)] = int(rng.choice([-1, 1])) * 30000
        else:
            coef = [0] * 64
            for i in rng.choice(64, int(rng.integers(1, 10)), replace=False):
                coef[int(i)] = int(rng.integers(-400, 401))
        levels[b] = [coef[ZIGZAG[p]] for p in range(64)]
    d = wl.DeviceUnit(ctx, u)
    ctx.wait(ctx.submit_mjpeg([d.struct]))
    out = u.host_outputs()
    assert oracle.oracle_mjpeg_reconstruct(C.byref(u.host_struct(out))) == 0
    assert np.array_equal(d.download("plane0"), out["plane0"])
    d.free()


def test_dv_fast_transform_fallback_bit_exact(ctx, oracle):
    """DV: the committed blocks of tests/golden/dv_fallback_blocks.json, on which the fused
    transform without its window rounds a sample differently from the reference (checked on the
    CPU by tests/test_idct_fast_path.py), placed among seeded blocks of both modes so that the
    kernel's mode-sorted order moves them between groups and waves, in a 625/50 and a 525/60 frame
    (the 525/60 one with every third block in 2-4-8 mode), against the oracle bit for bit."""
    fb = json.load(open(os.path.join(HERE, "golden", "dv_fallback_blocks.json")))
    special = fb["dv88"] + fb["dv248"]
    units = []
    for i, system in enumerate((abi.DV_SYSTEM_625_50, abi.DV_SYSTEM_525_60)):
        u = wl.dv_frame(seed=41, index=i, system=system, mode248_fraction=0.3 if i == 0 else 1.0 / 3.0)
        levels, info = u.inputs["levels"], u.inputs["info"]
        # fixtures at blocks 3, 10, 17, ...: spread over several 32-block tiles, mixed with both modes
        for k, b in enumerate(special * 3):
            pos = 3 + 7 * k + i
            levels[pos] = 0
            for p, v in b["levels"].items():
                levels[pos, int(p)] = v
            info[pos] = b["info"]
        units.append(u)
    dev = [wl.DeviceUnit(ctx, u) for u in units]
    ctx.wait(ctx.submit_dv([d.struct for d in dev]))
    for i, (u, d) in enumerate(zip(units, dev)):
        out = u.host_outputs()
        assert oracle.oracle_dv_reconstruct(C.byref(u.host_struct(out))) == 0
        for n in ("y", "cb", "cr"):
            assert np.array_equal(d.download(n), out[n]), (i, n)
        d.free()
