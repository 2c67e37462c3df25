"""Pins of the oracle's H.264 intra prediction.

* The reference's known answers: intra8_top_right_samples_are_available_for_the_top_right_partition
  (crates/codecs/h264/src/decoder.rs:15041-15055) and the five one-macroblock Intra16x16 slices of
  :17151-17330 (bitstreams :18613-18725), decoded here by hand into level records: QP 26
  (pic_init_qp 26, slice_qp_delta 0), mb_type 3 / 3 / 15 / 7 / 11 = I_16x16 with DC prediction
  and (chroma, luma) coded block patterns (0, 0) / (0, 0) / (0, 15) / (1, 0) / (2, 0), chroma DC
  prediction, deblocking off (disable_deblocking_filter_idc 1), one coefficient +2 at: nowhere /
  luma DC scan 0 / luma block 0 AC scan 1 / Cb DC 0 / Cb block 0 AC scan 1.
* A textbook restatement of ITU-T H.264 (05/2003+) 8.3.1.2 (Intra_4x4), 8.3.2.2 with the
  reference sample filtering 8.3.2.2.1 (Intra_8x8), 8.3.3 (Intra_16x16) and 8.3.4 (chroma, 4:2:0),
  written from the equations with the p[x, y] neighbour notation, against the oracle on random
  neighbours, every mode and every availability case. The reference's availability model is
  used: the corner sample p[-1, -1] exists exactly when the top and left neighbours do.
* Invariants: constant neighbours predict the constant in every mode.
"""
import ctypes as C

import numpy as np
import pytest

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl


# ------------------------------------------------------------------ textbook restatement

def _p_getter(top, left, corner):
    def p(x, y):
        if y == -1 and x == -1:
            return int(corner)
        if y == -1:
            return int(top[x])
        assert x == -1
        return int(left[y])
    return p


def tb_intra4(mode, top, left, corner):
    """8.3.1.2.1-9; top = p[0..7, -1] (with the substitution of p[4..7, -1] already done)."""
    p = _p_getter(top, left, corner)
    out = np.zeros((4, 4), np.int64)
    for y in range(4):
        for x in range(4):
            if mode == 0:
                v = p(x, -1)
            elif mode == 1:
                v = p(-1, y)
            elif mode == 2:
                if top is not None and left is not None:
                    v = (sum(p(i, -1) for i in range(4)) + sum(p(-1, i) for i in range(4)) + 4) >> 3
                elif left is not None:
                    v = (sum(p(-1, i) for i in range(4)) + 2) >> 2
                elif top is not None:
                    v = (sum(p(i, -1) for i in range(4)) + 2) >> 2
                else:
                    v = 128
            elif mode == 3:
                v = (p(6, -1) + 3 * p(7, -1) + 2) >> 2 if x == 3 and y == 3 else \
                    (p(x + y, -1) + 2 * p(x + y + 1, -1) + p(x + y + 2, -1) + 2) >> 2
            elif mode == 4:
                if x > y:
                    v = (p(x - y - 2, -1) + 2 * p(x - y - 1, -1) + p(x - y, -1) + 2) >> 2
                elif x < y:
                    v = (p(-1, y - x - 2) + 2 * p(-1, y - x - 1) + p(-1, y - x) + 2) >> 2
                else:
                    v = (p(0, -1) + 2 * p(-1, -1) + p(-1, 0) + 2) >> 2
            elif mode == 5:
                z = 2 * x - y
                if z >= 0 and z % 2 == 0:
                    v = (p(x - (y >> 1) - 1, -1) + p(x - (y >> 1), -1) + 1) >> 1
                elif z >= 0:
                    v = (p(x - (y >> 1) - 2, -1) + 2 * p(x - (y >> 1) - 1, -1) + p(x - (y >> 1), -1) + 2) >> 2
                elif z == -1:
                    v = (p(-1, 0) + 2 * p(-1, -1) + p(0, -1) + 2) >> 2
                else:
                    v = (p(-1, y - 1) + 2 * p(-1, y - 2) + p(-1, y - 3) + 2) >> 2
            elif mode == 6:
                z = 2 * y - x
                if z >= 0 and z % 2 == 0:
                    v = (p(-1, y - (x >> 1) - 1) + p(-1, y - (x >> 1)) + 1) >> 1
                elif z >= 0:
                    v = (p(-1, y - (x >> 1) - 2) + 2 * p(-1, y - (x >> 1) - 1) + p(-1, y - (x >> 1)) + 2) >> 2
                elif z == -1:
                    v = (p(-1, 0) + 2 * p(-1, -1) + p(0, -1) + 2) >> 2
                else:
                    v = (p(x - 1, -1) + 2 * p(x - 2, -1) + p(x - 3, -1) + 2) >> 2
            elif mode == 7:
                if y % 2 == 0:
                    v = (p(x + (y >> 1), -1) + p(x + (y >> 1) + 1, -1) + 1) >> 1
                else:
                    v = (p(x + (y >> 1), -1) + 2 * p(x + (y >> 1) + 1, -1) + p(x + (y >> 1) + 2, -1) + 2) >> 2
            else:
                z = x + 2 * y
                if z < 5 and z % 2 == 0:
                    v = (p(-1, y + (x >> 1)) + p(-1, y + (x >> 1) + 1) + 1) >> 1
                elif z < 5:
                    v = (p(-1, y + (x >> 1)) + 2 * p(-1, y + (x >> 1) + 1) + p(-1, y + (x >> 1) + 2) + 2) >> 2
                elif z == 5:
                    v = (p(-1, 2) + 3 * p(-1, 3) + 2) >> 2
                else:
                    v = p(-1, 3)
            out[y, x] = v
    return out


def tb_filter8(top, left, corner):
    """8.3.2.2.1 with the corner present exactly when top and left are."""
    ft = fl = fc = None
    if top is not None:
        t = [int(v) for v in top]
        ft = [0] * 16
        ft[0] = (corner + 2 * t[0] + t[1] + 2) >> 2 if corner is not None else (3 * t[0] + t[1] + 2) >> 2
        for x in range(1, 15):
            ft[x] = (t[x - 1] + 2 * t[x] + t[x + 1] + 2) >> 2
        ft[15] = (t[14] + 3 * t[15] + 2) >> 2
    if left is not None:
        l_ = [int(v) for v in left]
        fl = [0] * 8
        fl[0] = (corner + 2 * l_[0] + l_[1] + 2) >> 2 if corner is not None else (3 * l_[0] + l_[1] + 2) >> 2
        for y in range(1, 7):
            fl[y] = (l_[y - 1] + 2 * l_[y] + l_[y + 1] + 2) >> 2
        fl[7] = (l_[6] + 3 * l_[7] + 2) >> 2
    if corner is not None:
        fc = (int(top[0]) + 2 * int(corner) + int(left[0]) + 2) >> 2
    return ft, fl, fc


def tb_intra8(mode, top, left, corner):
    """8.3.2.2.2-10 on the filtered samples p'."""
    ft, fl, fc = tb_filter8(top, left, corner)
    p = _p_getter(ft, fl, fc)
    out = np.zeros((8, 8), np.int64)
    for y in range(8):
        for x in range(8):
            if mode == 0:
                v = p(x, -1)
            elif mode == 1:
                v = p(-1, y)
            elif mode == 2:
                if ft is not None and fl is not None:
                    v = (sum(ft[:8]) + sum(fl) + 8) >> 4
                elif fl is not None:
                    v = (sum(fl) + 4) >> 3
                elif ft is not None:
                    v = (sum(ft[:8]) + 4) >> 3
                else:
                    v = 128
            elif mode == 3:
                v = (p(14, -1) + 3 * p(15, -1) + 2) >> 2 if x == 7 and y == 7 else \
                    (p(x + y, -1) + 2 * p(x + y + 1, -1) + p(x + y + 2, -1) + 2) >> 2
            elif mode == 4:
                if x > y:
                    v = (p(x - y - 2, -1) + 2 * p(x - y - 1, -1) + p(x - y, -1) + 2) >> 2
                elif x < y:
                    v = (p(-1, y - x - 2) + 2 * p(-1, y - x - 1) + p(-1, y - x) + 2) >> 2
                else:
                    v = (p(0, -1) + 2 * p(-1, -1) + p(-1, 0) + 2) >> 2
            elif mode == 5:
                z = 2 * x - y
                if z >= 0 and z % 2 == 0:
                    v = (p(x - (y >> 1) - 1, -1) + p(x - (y >> 1), -1) + 1) >> 1
                elif z >= 0:
                    v = (p(x - (y >> 1) - 2, -1) + 2 * p(x - (y >> 1) - 1, -1) + p(x - (y >> 1), -1) + 2) >> 2
                elif z == -1:
                    v = (p(-1, 0) + 2 * p(-1, -1) + p(0, -1) + 2) >> 2
                else:
                    v = (p(-1, y - 2 * x - 1) + 2 * p(-1, y - 2 * x - 2) + p(-1, y - 2 * x - 3) + 2) >> 2
            elif mode == 6:
                z = 2 * y - x
                if z >= 0 and z % 2 == 0:
                    v = (p(-1, y - (x >> 1) - 1) + p(-1, y - (x >> 1)) + 1) >> 1
                elif z >= 0:
                    v = (p(-1, y - (x >> 1) - 2) + 2 * p(-1, y - (x >> 1) - 1) + p(-1, y - (x >> 1)) + 2) >> 2
                elif z == -1:
                    v = (p(-1, 0) + 2 * p(-1, -1) + p(0, -1) + 2) >> 2
                else:
                    v = (p(x - 2 * y - 1, -1) + 2 * p(x - 2 * y - 2, -1) + p(x - 2 * y - 3, -1) + 2) >> 2
            elif mode == 7:
                if y % 2 == 0:
                    v = (p(x + (y >> 1), -1) + p(x + (y >> 1) + 1, -1) + 1) >> 1
                else:
                    v = (p(x + (y >> 1), -1) + 2 * p(x + (y >> 1) + 1, -1) + p(x + (y >> 1) + 2, -1) + 2) >> 2
            else:
                z = x + 2 * y
                if z < 13 and z % 2 == 0:
                    v = (p(-1, y + (x >> 1)) + p(-1, y + (x >> 1) + 1) + 1) >> 1
                elif z < 13:
                    v = (p(-1, y + (x >> 1)) + 2 * p(-1, y + (x >> 1) + 1) + p(-1, y + (x >> 1) + 2) + 2) >> 2
                elif z == 13:
                    v = (p(-1, 6) + 3 * p(-1, 7) + 2) >> 2
                else:
                    v = p(-1, 7)
            out[y, x] = v
    return out


def tb_plane(top, left, corner, size):
    """8.3.3.4 (size 16) and 8.3.4.4 (size 8, 4:2:0: b = (34 H + 32) >> 6)."""
    n = size // 2
    p = _p_getter(top, left, corner)
    H = sum((i + 1) * (p(n + i, -1) - p(n - 2 - i, -1)) for i in range(n))
    V = sum((i + 1) * (p(-1, n + i) - p(-1, n - 2 - i)) for i in range(n))
    a = 16 * (p(-1, size - 1) + p(size - 1, -1))
    if size == 16:
        b, c = (5 * H + 32) >> 6, (5 * V + 32) >> 6
    else:
        b, c = (34 * H + 32) >> 6, (34 * V + 32) >> 6
    out = np.zeros((size, size), np.int64)
    for y in range(size):
        for x in range(size):
            out[y, x] = min(255, max(0, (a + b * (x - (n - 1)) + c * (y - (n - 1)) + 16) >> 5))
    return out


def tb_intra16(mode, top, left, corner):
    if mode == 0:
        return np.tile(np.asarray(top, np.int64), (16, 1))
    if mode == 1:
        return np.tile(np.asarray(left, np.int64)[:, None], (1, 16))
    if mode == 2:
        if top is not None and left is not None:
            v = (int(np.sum(top)) + int(np.sum(left)) + 16) >> 5
        elif left is not None:
            v = (int(np.sum(left)) + 8) >> 4
        elif top is not None:
            v = (int(np.sum(top)) + 8) >> 4
        else:
            v = 128
        return np.full((16, 16), v, np.int64)
    return tb_plane(top, left, corner, 16)


def tb_chroma(mode, top, left, corner):
    """8.3.4: mode 0 DC per 4x4 chroma block, 1 horizontal, 2 vertical, 3 plane."""
    if mode == 1:
        return np.tile(np.asarray(left, np.int64)[:, None], (1, 8))
    if mode == 2:
        return np.tile(np.asarray(top, np.int64), (8, 1))
    if mode == 3:
        return tb_plane(top, left, corner, 8)
    out = np.zeros((8, 8), np.int64)
    for yo in (0, 4):
        for xo in (0, 4):
            s_t = int(np.sum(top[xo:xo + 4])) if top is not None else None
            s_l = int(np.sum(left[yo:yo + 4])) if left is not None else None
            if (xo, yo) in ((0, 0), (4, 4)):
                if s_t is not None and s_l is not None:
                    v = (s_t + s_l + 4) >> 3
                elif s_l is not None:
                    v = (s_l + 2) >> 2
                elif s_t is not None:
                    v = (s_t + 2) >> 2
                else:
                    v = 128
            elif (xo, yo) == (4, 0):
                v = (s_t + 2) >> 2 if s_t is not None else ((s_l + 2) >> 2 if s_l is not None else 128)
            else:
                v = (s_l + 2) >> 2 if s_l is not None else ((s_t + 2) >> 2 if s_t is not None else 128)
            out[yo:yo + 4, xo:xo + 4] = v
    return out


# ------------------------------------------------------------------ oracle calls on a plane

def _plane_with(rng, size_w, size_h):
    return rng.integers(0, 256, (size_h, size_w), dtype=np.uint8)


NEEDS = {0: "t", 1: "l", 2: "", 3: "t", 4: "tlc", 5: "tlc", 6: "tlc", 7: "t", 8: "l"}


@pytest.mark.parametrize("mode", range(9))
def test_intra4_matches_textbook(oracle, mode):
    rng = np.random.default_rng(40 + mode)
    W = 48
    for trial in range(200):
        plane = _plane_with(rng, W, 48)
        # a 4x4 block inside macroblock (1, 1): block index k, availability of the macroblocks
        k = int(rng.integers(0, 16))
        top_mb, left_mb = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        bx, by = wl.BLOCK_XY[k]
        ox, oy = 16 + 4 * bx, 16 + 4 * by
        has_t, has_l = by > 0 or top_mb, bx > 0 or left_mb
        if ("t" in NEEDS[mode] and not has_t) or ("l" in NEEDS[mode] and not has_l):
            out = np.zeros(16, np.uint8)
            assert oracle.oracle_h264_intra4_predict(plane.ctypes.data, W, ox, oy, k, mode, top_mb, left_mb,
                                                     out.ctypes.data) == -1
            continue
        top = left = corner = None
        if has_t:
            top = [int(plane[oy - 1, ox + x]) for x in range(4)]
            # 8.3.1.2: p[4..7, -1] from the block to the upper right when it is already decoded and
            # inside the picture, else copies of p[3, -1] (the reference's rule set)
            tr_ok = k not in (3, 11) and not (oy % 16 != 0 and ox % 16 + 4 >= 16) and ox + 4 < W
            top += [int(plane[oy - 1, ox + x]) if tr_ok else top[3] for x in range(4, 8)]
        if has_l:
            left = [int(plane[oy + y, ox - 1]) for y in range(4)]
        if has_t and has_l:
            corner = int(plane[oy - 1, ox - 1])
        out = np.zeros(16, np.uint8)
        assert oracle.oracle_h264_intra4_predict(plane.ctypes.data, W, ox, oy, k, mode, top_mb, left_mb,
                                                 out.ctypes.data) == 0
        np.testing.assert_array_equal(out.reshape(4, 4), tb_intra4(mode, top, left, corner), err_msg=f"{trial} k={k}")


@pytest.mark.parametrize("mode", range(9))
def test_intra8_matches_textbook(oracle, mode):
    rng = np.random.default_rng(60 + mode)
    W = 48
    for trial in range(120):
        plane = _plane_with(rng, W, 48)
        g = int(rng.integers(0, 4))
        top_mb, left_mb = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        mbx = int(rng.integers(1, 3))   # macroblock column 1 or 2 (2: the top-right of group 1 is past the edge)
        ox, oy = 16 * mbx + 8 * (g % 2), 16 + 8 * (g // 2)
        has_t, has_l = g >= 2 or top_mb, g % 2 == 1 or left_mb
        out = np.zeros(64, np.uint8)
        rc = oracle.oracle_h264_intra8_predict(plane.ctypes.data, W, ox, oy, g, mode, top_mb, left_mb, out.ctypes.data)
        if ("t" in NEEDS[mode] and not has_t) or ("l" in NEEDS[mode] and not has_l):
            assert rc == -1
            continue
        assert rc == 0
        top = left = corner = None
        if has_t:
            top = [int(plane[oy - 1, ox + x]) for x in range(8)]
            tr_ok = g != 3 and ox + 8 < W
            top += [int(plane[oy - 1, ox + x]) if tr_ok else top[7] for x in range(8, 16)]
        if has_l:
            left = [int(plane[oy + y, ox - 1]) for y in range(8)]
        if has_t and has_l:
            corner = int(plane[oy - 1, ox - 1])
        np.testing.assert_array_equal(out.reshape(8, 8), tb_intra8(mode, top, left, corner), err_msg=f"{trial} g={g}")


@pytest.mark.parametrize("mode", range(4))
def test_intra16_and_chroma_match_textbook(oracle, mode):
    rng = np.random.default_rng(80 + mode)
    for trial in range(150):
        plane = _plane_with(rng, 48, 48)
        has_t, has_l = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        top = [int(v) for v in plane[15, 16:32]] if has_t else None
        left = [int(v) for v in plane[16:32, 15]] if has_l else None
        corner = int(plane[15, 15]) if has_t and has_l else None
        out = np.zeros(256, np.uint8)
        rc = oracle.oracle_h264_intra_block_predict(plane.ctypes.data, 48, 16, 16, 16, mode, has_t, has_l, out.ctypes.data)
        needs = {0: (True, False), 1: (False, True), 2: (False, False), 3: (True, True)}[mode]
        if (needs[0] and not has_t) or (needs[1] and not has_l):
            assert rc == -1
        else:
            assert rc == 0
            np.testing.assert_array_equal(out.reshape(16, 16), tb_intra16(mode, top, left, corner), err_msg=str(trial))
        # chroma 8x8 at (8, 8) of a 24-wide plane
        cp = plane[:24, :24].copy()
        ctop = [int(v) for v in cp[7, 8:16]] if has_t else None
        cleft = [int(v) for v in cp[8:16, 7]] if has_l else None
        ccorner = int(cp[7, 7]) if has_t and has_l else None
        cout = np.zeros(64, np.uint8)
        rc = oracle.oracle_h264_chroma_predict(cp.ctypes.data, 24, 8, 8, mode, has_t, has_l, cout.ctypes.data)
        cneeds = {0: (False, False), 1: (False, True), 2: (True, False), 3: (True, True)}[mode]
        if (cneeds[0] and not has_t) or (cneeds[1] and not has_l):
            assert rc == -1
        else:
            assert rc == 0
            np.testing.assert_array_equal(cout.reshape(8, 8), tb_chroma(mode, ctop, cleft, ccorner), err_msg=str(trial))


def test_constant_neighbours_predict_the_constant(oracle):
    for v in (0, 77, 255):
        plane = np.full((48, 48), v, np.uint8)
        out4, out8, out16, outc = (np.zeros(n, np.uint8) for n in (16, 64, 256, 64))
        for mode in range(9):
            assert oracle.oracle_h264_intra4_predict(plane.ctypes.data, 48, 20, 20, 3, mode, 1, 1, out4.ctypes.data) == 0
            assert (out4 == v).all(), (v, mode)
            assert oracle.oracle_h264_intra8_predict(plane.ctypes.data, 48, 16, 16, 0, mode, 1, 1, out8.ctypes.data) == 0
            assert (out8 == v).all(), (v, mode)
        for mode in range(4):
            assert oracle.oracle_h264_intra_block_predict(plane.ctypes.data, 48, 16, 16, 16, mode, 1, 1, out16.ctypes.data) == 0
            assert (out16 == v).all()
            assert oracle.oracle_h264_chroma_predict(plane.ctypes.data, 48, 16, 16, mode, 1, 1, outc.ctypes.data) == 0
            assert (outc == v).all()


# ------------------------------------------------------------------ the reference's known answers

def test_intra8_top_right_reference_kat(oracle):
    stride = 32
    plane = np.zeros((32, stride), np.uint8)
    plane[7, 14:17] = (40, 80, 120)
    out = np.zeros(64, np.uint8)
    assert oracle.oracle_h264_intra8_predict(plane.ctypes.data, stride, 8, 8, 1, 0, 1, 1, out.ctypes.data) == 0
    assert out[7] == 80
    plane[15, 14:17] = (40, 80, 120)
    assert oracle.oracle_h264_intra8_predict(plane.ctypes.data, stride, 8, 16, 3, 0, 1, 1, out.ctypes.data) == 0
    assert out[7] == 70


def intra16_kat_picture(case):
    """One 16x16 Intra16x16 (DC prediction) macroblock, QP 26, as the reference's slices code it."""
    mbs = np.zeros(1, wl.MB_DTYPE)
    mbs["flags"] = abi.H264_MB_INTRA
    mbs["qp"] = 26
    levels = np.zeros((1, abi.H264_LEVELS_PER_MB), np.int16)
    luma_cbp, chroma_cbp = {"zero": (0, 0), "luma_dc": (0, 0), "luma_ac": (15, 0), "chroma_dc": (0, 1),
                            "chroma_ac": (0, 2)}[case]
    if case == "luma_dc":
        levels[0, 0] = 2           # DC scan index 0
    elif case == "luma_ac":
        levels[0, 1] = 2           # block 0, AC scan index 1
    elif case == "chroma_dc":
        levels[0, 256] = 2         # Cb DC c0
    elif case == "chroma_ac":
        levels[0, 264] = 2         # Cb block 0, AC scan index 1
    mbs["cbp"] = luma_cbp | chroma_cbp << 4
    intra = np.zeros(1, wl.INTRA_DTYPE)
    intra["type"] = abi.H264_INTRA_16X16
    intra["luma_mode"] = 2
    intra["chroma_mode"] = 0
    blocks = np.zeros((1, 16), wl.BLOCK_DTYPE)
    blocks["ref_idx"] = -1
    return mbs, blocks, levels, intra


KAT_EXPECT = {
    # name: (luma check, cb check, cr check); a tuple is the first four samples of the plane
    "zero": (128, 128, 128),
    "luma_dc": (130, 128, 128),
    "luma_ac": ((136, 132, 124, 120), None, None),
    "chroma_dc": (None, 131, None),
    "chroma_ac": (None, (136, 132, 124, 120), None),
}


def run_oracle_picture(oracle, mbs, blocks, levels, intra, wmb=1, hmb=1):
    y = np.zeros((16 * hmb, 16 * wmb), np.uint8)
    cb = np.zeros((8 * hmb, 8 * wmb), np.uint8)
    cr = np.zeros((8 * hmb, 8 * wmb), np.uint8)
    pic = abi.new(abi.H264Picture, width_mbs=wmb, height_mbs=hmb, deblock=0, macroblocks=mbs.ctypes.data,
                  blocks=blocks.ctypes.data, levels=levels.ctypes.data, intra=intra.ctypes.data)
    pic.planes[0], pic.planes[1], pic.planes[2] = y.ctypes.data, cb.ctypes.data, cr.ctypes.data
    assert oracle.oracle_h264_reconstruct(C.byref(pic)) == 0
    return {"y": y, "cb": cb, "cr": cr}


@pytest.mark.parametrize("case", sorted(KAT_EXPECT))
def test_intra16_slices_reference_kats(oracle, case):
    out = run_oracle_picture(oracle, *intra16_kat_picture(case))
    for plane, want in zip(("y", "cb", "cr"), KAT_EXPECT[case]):
        if want is None:
            continue
        if isinstance(want, tuple):
            assert tuple(out[plane].reshape(-1)[:4]) == want, (case, plane)
        else:
            assert (out[plane] == want).all(), (case, plane)


def test_intra_without_records_is_unsupported(oracle):
    mbs, blocks, levels, intra = intra16_kat_picture("zero")
    y = np.zeros((16, 16), np.uint8)
    pic = abi.new(abi.H264Picture, width_mbs=1, height_mbs=1, deblock=0, macroblocks=mbs.ctypes.data,
                  blocks=blocks.ctypes.data, levels=levels.ctypes.data)
    pic.planes[0] = pic.planes[1] = pic.planes[2] = y.ctypes.data
    assert oracle.oracle_h264_reconstruct(C.byref(pic)) == -3


def test_intra16_dc_transform_matches_textbook(oracle):
    """8.5.10: c = the 4x4 DC array (inverse zig-zag), f = H c H with H the 4x4 Hadamard matrix,
    dcY = (f LevelScale4x4(qP % 6, 0, 0)) << (qP / 6 - 6) for qP >= 36, else
    (f LevelScale + 2^(5 - qP / 6)) >> (6 - qP / 6); LevelScale = 16 * normAdjust (flat list)."""
    zz = [(0, 0), (0, 1), (1, 0), (2, 0), (1, 1), (0, 2), (0, 3), (1, 2), (2, 1), (3, 0), (3, 1), (2, 2), (1, 3),
          (2, 3), (3, 2), (3, 3)]
    H = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]], np.int64)
    norm = [10, 11, 13, 14, 16, 18]
    flat = np.full(16, 16, np.uint8)
    rng = np.random.default_rng(3)
    for qp in range(52):
        for _ in range(10):
            lv = rng.integers(-300, 301, 16).astype(np.int32)
            c = np.zeros((4, 4), np.int64)
            for i, (r, k) in enumerate(zz):
                c[r, k] = lv[i]
            f = H @ c @ H
            ls = 16 * norm[qp % 6]
            want = (f * ls) << (qp // 6 - 6) if qp >= 36 else (f * ls + (1 << (5 - qp // 6))) >> (6 - qp // 6)
            out = np.zeros(16, np.int32)
            assert oracle.oracle_h264_intra16_dc(lv.ctypes.data, qp, flat.ctypes.data, out.ctypes.data) == 0
            np.testing.assert_array_equal(out.reshape(4, 4), want, err_msg=str(qp))
