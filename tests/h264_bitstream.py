"""H.264 CAVLC bitstream writer for the front-end tests (tests/test_h264_entropy.py,
tests/test_gpu_h264_streams.py): parameter sets, slice headers and macroblock layers written from
ITU-T H.264 7.3 (syntax), 9.1 (Exp-Golomb), 9.2 (CAVLC residual) and Table 9-4 (coded block
pattern), with the encoder-side context the syntax needs (nC of each residual block, 9.2.1; the
Intra4x4 / Intra8x8 mode predictor, 8.3.1.1) kept independently of the decoder.

A stream is built as access units (lists of NAL units); `length_prefixed` packs one access unit as
the front end takes it (4-byte big-endian lengths)."""
import numpy as np

# ---------------------------------------------------------------- CAVLC code tables (Tables 9-5 .. 9-10)
# coeff_token (length, code) per nC class, index 4 * TotalCoeff + TrailingOnes; chroma DC nC = -1.
COEFF_TOKEN = [
    ([1, 0, 0, 0, 6, 2, 0, 0, 8, 6, 3, 0, 9, 8, 7, 5, 10, 9, 8, 6, 11, 10, 9, 7, 13, 11, 10, 8, 13, 13, 11, 9, 13, 13, 13,
      10, 14, 14, 13, 11, 14, 14, 14, 13, 15, 15, 14, 14, 15, 15, 15, 14, 16, 15, 15, 15, 16, 16, 16, 15, 16, 16, 16, 16,
      16, 16, 16, 16],
     [1, 0, 0, 0, 5, 1, 0, 0, 7, 4, 1, 0, 7, 6, 5, 3, 7, 6, 5, 3, 7, 6, 5, 4, 15, 6, 5, 4, 11, 14, 5, 4, 8, 10, 13, 4, 15,
      14, 9, 4, 11, 10, 13, 12, 15, 14, 9, 12, 11, 10, 13, 8, 15, 1, 9, 12, 11, 14, 13, 8, 7, 10, 9, 12, 4, 6, 5, 8]),
    ([2, 0, 0, 0, 6, 2, 0, 0, 6, 5, 3, 0, 7, 6, 6, 4, 8, 6, 6, 4, 8, 7, 7, 5, 9, 8, 8, 6, 11, 9, 9, 6, 11, 11, 11, 7, 12,
      11, 11, 9, 12, 12, 12, 11, 12, 12, 12, 11, 13, 13, 13, 12, 13, 13, 13, 13, 13, 14, 13, 13, 14, 14, 14, 13, 14, 14,
      14, 14],
     [3, 0, 0, 0, 11, 2, 0, 0, 7, 7, 3, 0, 7, 10, 9, 5, 7, 6, 5, 4, 4, 6, 5, 6, 7, 6, 5, 8, 15, 6, 5, 4, 11, 14, 13, 4,
      15, 10, 9, 4, 11, 14, 13, 12, 8, 10, 9, 8, 15, 14, 13, 12, 11, 10, 9, 12, 7, 11, 6, 8, 9, 8, 10, 1, 7, 6, 5, 4]),
    ([4, 0, 0, 0, 6, 4, 0, 0, 6, 5, 4, 0, 6, 5, 5, 4, 7, 5, 5, 4, 7, 5, 5, 4, 7, 6, 6, 4, 7, 6, 6, 4, 8, 7, 7, 5, 8, 8,
      7, 6, 9, 8, 8, 7, 9, 9, 8, 8, 9, 9, 9, 8, 10, 9, 9, 9, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10],
     [15, 0, 0, 0, 15, 14, 0, 0, 11, 15, 13, 0, 8, 12, 14, 12, 15, 10, 11, 11, 11, 8, 9, 10, 9, 14, 13, 9, 8, 10, 9, 8,
      15, 14, 13, 13, 11, 14, 10, 12, 15, 10, 13, 12, 11, 14, 9, 12, 8, 10, 13, 8, 13, 7, 9, 12, 9, 12, 11, 10, 5, 8, 7,
      6, 1, 4, 3, 2]),
    ([6, 0, 0, 0, 6, 6, 0, 0, 6, 6, 6, 0] + [6] * 56,
     [3, 0, 0, 0, 0, 1, 0, 0, 4, 5, 6, 0] + list(range(8, 64))),
]
CHROMA_DC_TOKEN = ([2, 0, 0, 0, 6, 1, 0, 0, 6, 6, 3, 0, 6, 7, 7, 6, 6, 8, 8, 7],
                   [1, 0, 0, 0, 7, 1, 0, 0, 4, 6, 1, 0, 3, 3, 2, 5, 2, 3, 2, 0])
TOTAL_ZEROS = [  # [TotalCoeff - 1] -> (lengths, codes) by total_zeros (4x4 blocks)
    ([1, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 9], [1, 3, 2, 3, 2, 3, 2, 3, 2, 3, 2, 3, 2, 3, 2, 1]),
    ([3, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 6, 6, 6, 6], [7, 6, 5, 4, 3, 5, 4, 3, 2, 3, 2, 3, 2, 1, 0]),
    ([4, 3, 3, 3, 4, 4, 3, 3, 4, 5, 5, 6, 5, 6], [5, 7, 6, 5, 4, 3, 4, 3, 2, 3, 2, 1, 1, 0]),
    ([5, 3, 4, 4, 3, 3, 3, 4, 3, 4, 5, 5, 5], [3, 7, 5, 4, 6, 5, 4, 3, 3, 2, 2, 1, 0]),
    ([4, 4, 4, 3, 3, 3, 3, 3, 4, 5, 4, 5], [5, 4, 3, 7, 6, 5, 4, 3, 2, 1, 1, 0]),
    ([6, 5, 3, 3, 3, 3, 3, 3, 4, 3, 6], [1, 1, 7, 6, 5, 4, 3, 2, 1, 1, 0]),
    ([6, 5, 3, 3, 3, 2, 3, 4, 3, 6], [1, 1, 5, 4, 3, 3, 2, 1, 1, 0]),
    ([6, 4, 5, 3, 2, 2, 3, 3, 6], [1, 1, 1, 3, 3, 2, 2, 1, 0]),
    ([6, 6, 4, 2, 2, 3, 2, 5], [1, 0, 1, 3, 2, 1, 1, 1]),
    ([5, 5, 3, 2, 2, 2, 4], [1, 0, 1, 3, 2, 1, 1]),
    ([4, 4, 3, 3, 1, 3], [0, 1, 1, 2, 1, 3]),
    ([4, 4, 2, 1, 3], [0, 1, 1, 1, 1]),
    ([3, 3, 1, 2], [0, 1, 1, 1]),
    ([2, 2, 1], [0, 1, 1]),
    ([1, 1], [0, 1]),
]
CHROMA_DC_TOTAL_ZEROS = [([1, 2, 3, 3], [1, 1, 1, 0]), ([1, 2, 2], [1, 1, 0]), ([1, 1], [1, 0])]
RUN_BEFORE = [  # [min(zerosLeft, 7) - 1] -> (lengths, codes) by run_before
    ([1, 1], [1, 0]),
    ([1, 2, 2], [1, 1, 0]),
    ([2, 2, 2, 2], [3, 2, 1, 0]),
    ([2, 2, 2, 3, 3], [3, 2, 1, 1, 0]),
    ([2, 2, 3, 3, 3, 3], [3, 2, 3, 2, 1, 0]),
    ([2, 3, 3, 3, 3, 3, 3], [3, 0, 1, 3, 2, 5, 4]),
    ([3, 3, 3, 3, 3, 3, 3, 4, 5, 6, 7, 8, 9, 10, 11], [7, 6, 5, 4, 3, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1]),
]
# Table 9-4 (ChromaArrayType 1): codeNum -> coded_block_pattern for Intra_4x4 / Intra_8x8 and Inter.
CBP_INTRA = [47, 31, 15, 0, 23, 27, 29, 30, 7, 11, 13, 14, 39, 43, 45, 46, 16, 3, 5, 10, 12, 19, 21, 26, 28, 35, 37, 42,
             44, 1, 2, 4, 8, 17, 18, 20, 24, 6, 9, 22, 25, 32, 33, 34, 36, 40, 38, 41]
CBP_INTER = [0, 16, 1, 2, 4, 8, 32, 3, 5, 10, 12, 15, 47, 7, 11, 13, 14, 6, 9, 31, 35, 37, 42, 44, 33, 34, 36, 40, 39, 43,
             45, 46, 17, 18, 20, 24, 19, 21, 26, 28, 23, 27, 29, 30, 22, 25, 38, 41]


def prefix_code_ok(lengths, codes):
    """The (length, code) pairs with length > 0 form a prefix-free code whose Kraft sum is <= 1."""
    words = [format(c, f"0{n}b") for n, c in zip(lengths, codes) if n]
    if any(len(w) != n for w, n in zip(words, [n for n in lengths if n])):
        return False, 0.0
    for i, a in enumerate(words):
        for j, b in enumerate(words):
            if i != j and b.startswith(a):
                return False, 0.0
    return True, sum(2.0 ** -len(w) for w in words)


# ---------------------------------------------------------------- bits
class BitWriter:
    def __init__(self):
        self.bits = []

    def u(self, value, n):
        for i in range(n - 1, -1, -1):
            self.bits.append((int(value) >> i) & 1)

    def bit(self, b):
        self.bits.append(1 if b else 0)

    def ue(self, v):
        v = int(v) + 1
        n = v.bit_length()
        self.u(0, n - 1)
        self.u(v, n)

    def se(self, v):
        v = int(v)
        self.ue(2 * v - 1 if v > 0 else -2 * v)

    def te(self, v, rng):
        if rng == 1:
            self.bit(1 - v)
        else:
            self.ue(v)

    def align_zero(self):
        while len(self.bits) % 8:
            self.bits.append(0)

    def rbsp(self):
        bits = self.bits + [1]
        while len(bits) % 8:
            bits.append(0)
        return bytes(int("".join(map(str, bits[i:i + 8])), 2) for i in range(0, len(bits), 8))


def nal(nal_ref_idc, nal_type, rbsp):
    """NAL unit: header byte, then the RBSP with emulation prevention (7.4.1.1)."""
    out = bytearray([(nal_ref_idc << 5) | nal_type])
    zeros = 0
    for byte in rbsp:
        if zeros >= 2 and byte <= 3:
            out.append(3)
            zeros = 0
        out.append(byte)
        zeros = zeros + 1 if byte == 0 else 0
    return bytes(out)


def length_prefixed(nals):
    return b"".join(len(n).to_bytes(4, "big") + n for n in nals)


# ---------------------------------------------------------------- parameter sets and slice headers
def sps(width_mbs=1, height_mbs=1, max_refs=2, profile=66, poc_type=0, log2_max_frame_num=4, log2_max_poc_lsb=4,
        direct_8x8_inference=True, transform_bypass=False):
    w = BitWriter()
    w.u(profile, 8)
    w.u(0, 8)
    w.u(10 if profile == 66 else 40, 8)
    w.ue(0)
    if profile in (100, 110, 122, 244, 44, 83, 86, 118, 128, 138, 139, 134, 135):
        w.ue(1)        # chroma_format_idc 4:2:0
        w.ue(0)
        w.ue(0)
        w.bit(transform_bypass)
        w.bit(0)       # no scaling matrices
    w.ue(log2_max_frame_num - 4)
    w.ue(poc_type)
    if poc_type == 0:
        w.ue(log2_max_poc_lsb - 4)
    w.ue(max_refs)
    w.bit(0)
    w.ue(width_mbs - 1)
    w.ue(height_mbs - 1)
    w.bit(1)           # frame_mbs_only
    w.bit(direct_8x8_inference)
    w.bit(0)           # no cropping
    w.bit(0)           # no VUI
    return nal(3, 7, w.rbsp())


def pps(weighted_pred=False, weighted_bipred_idc=0, init_qp=26, chroma_qp_offset=0, deblocking_control=True,
        transform_8x8=False, second_chroma_qp_offset=None, num_ref_idx_default=(1, 1)):
    w = BitWriter()
    w.ue(0)
    w.ue(0)
    w.bit(0)           # CAVLC
    w.bit(0)
    w.ue(0)
    w.ue(num_ref_idx_default[0] - 1)
    w.ue(num_ref_idx_default[1] - 1)
    w.bit(weighted_pred)
    w.u(weighted_bipred_idc, 2)
    w.se(init_qp - 26)
    w.se(0)
    w.se(chroma_qp_offset)
    w.bit(deblocking_control)
    w.bit(0)
    w.bit(0)
    if transform_8x8 or second_chroma_qp_offset is not None:
        w.bit(transform_8x8)
        w.bit(0)
        w.se(chroma_qp_offset if second_chroma_qp_offset is None else second_chroma_qp_offset)
    return nal(3, 8, w.rbsp())


class Params:
    """Stream-level values the slice syntax depends on."""

    def __init__(self, width_mbs=1, height_mbs=1, log2_max_frame_num=4, log2_max_poc_lsb=4, weighted_pred=False,
                 weighted_bipred_idc=0, init_qp=26, deblocking_control=True, transform_8x8=False,
                 num_ref_idx_default=(1, 1)):
        self.width_mbs, self.height_mbs = width_mbs, height_mbs
        self.log2_max_frame_num, self.log2_max_poc_lsb = log2_max_frame_num, log2_max_poc_lsb
        self.weighted_pred, self.weighted_bipred_idc = weighted_pred, weighted_bipred_idc
        self.init_qp, self.deblocking_control, self.transform_8x8 = init_qp, deblocking_control, transform_8x8
        self.num_ref_idx_default = num_ref_idx_default


def slice_header(w, prm, slice_type, frame_num, poc_lsb, idr=False, nal_ref_idc=1, first_mb=0, num_ref_idx=None,
                 direct_spatial=True, weights=None, qp=None, disable_deblocking=1, alpha=0, beta=0,
                 long_term_reference=False, modifications=None):
    """slice_type 0 P, 1 B, 2 I. weights: {"luma_denom", "chroma_denom", "l0": [(lw, lo, [(cw, co), (cw, co)]) or None, ...],
    "l1": ...}; modifications: {list: [(idc, value), ...]}."""
    w.ue(first_mb)
    w.ue(slice_type)
    w.ue(0)
    w.u(frame_num, prm.log2_max_frame_num)
    if idr:
        w.ue(0)
    w.u(poc_lsb, prm.log2_max_poc_lsb)
    if slice_type == 1:
        w.bit(direct_spatial)
    active = list(prm.num_ref_idx_default)
    if slice_type != 2:
        if num_ref_idx is not None:
            w.bit(1)
            w.ue(num_ref_idx[0] - 1)
            if slice_type == 1:
                w.ue(num_ref_idx[1] - 1)
            active = list(num_ref_idx)
        else:
            w.bit(0)
        for lst in range(2 if slice_type == 1 else 1):
            mods = (modifications or {}).get(lst)
            w.bit(bool(mods))
            if mods:
                for idc, value in mods:
                    w.ue(idc)
                    w.ue(value)
                w.ue(3)
    if (prm.weighted_pred and slice_type == 0) or (prm.weighted_bipred_idc == 1 and slice_type == 1):
        w.ue(weights["luma_denom"])
        w.ue(weights["chroma_denom"])
        for lst in range(2 if slice_type == 1 else 1):
            for i in range(active[lst]):
                entry = weights["l%d" % lst][i]
                if entry is None:
                    w.bit(0)
                    w.bit(0)
                    continue
                lw, lo, chroma = entry
                w.bit(1)
                w.se(lw)
                w.se(lo)
                w.bit(chroma is not None)
                if chroma is not None:
                    for cw, co in chroma:
                        w.se(cw)
                        w.se(co)
    if nal_ref_idc:
        if idr:
            w.bit(0)
            w.bit(long_term_reference)
        else:
            w.bit(0)   # sliding window
    w.se((prm.init_qp if qp is None else qp) - prm.init_qp)
    if prm.deblocking_control:
        w.ue(disable_deblocking)
        if disable_deblocking != 1:
            w.se(alpha // 2)
            w.se(beta // 2)
    return active


# ---------------------------------------------------------------- CAVLC residual encoding
def _write_vlc(w, lengths, codes, index):
    n = lengths[index]
    assert n, index
    w.u(codes[index], n)


def _write_level(w, level, suffix_length, first_non_trailing):
    code = 2 * abs(level) - (2 if level > 0 else 1) - (2 if first_non_trailing else 0)
    assert code >= 0
    for prefix in range(0, 32):
        size = 4 if (prefix == 14 and suffix_length == 0) else (prefix - 3 if prefix >= 15 else suffix_length)
        base = min(prefix, 15) << suffix_length
        if prefix >= 15 and suffix_length == 0:
            base += 15
        if prefix >= 16:
            base += (1 << (prefix - 3)) - 4096
        if base <= code < base + (1 << size):
            w.u(0, prefix)
            w.bit(1)
            if size:
                w.u(code - base, size)
            return
    raise ValueError("level too large")


def encode_block(w, nc, coeffs):
    """residual_block_cavlc for coefficients in scan order (len 4, 15 or 16). Returns TotalCoeff."""
    n = len(coeffs)
    nz = [(i, int(v)) for i, v in enumerate(coeffs) if v]
    total = len(nz)
    levels = [v for _, v in reversed(nz)]
    t1 = 0
    for v in levels[:3]:
        if abs(v) != 1:
            break
        t1 += 1
    if nc == -1:
        _write_vlc(w, *CHROMA_DC_TOKEN, 4 * total + t1)
    else:
        _write_vlc(w, *COEFF_TOKEN[0 if nc < 2 else 1 if nc < 4 else 2 if nc < 8 else 3], 4 * total + t1)
    if total == 0:
        return 0
    for v in levels[:t1]:
        w.bit(v < 0)
    suffix_length = 1 if total > 10 and t1 < 3 else 0
    for k, v in enumerate(levels[t1:]):
        _write_level(w, v, suffix_length, k == 0 and t1 < 3)
        if suffix_length == 0:
            suffix_length = 1
        if abs(v) > (3 << (suffix_length - 1)) and suffix_length < 6:
            suffix_length += 1
    total_zeros = nz[-1][0] + 1 - total
    if total < n:
        if n == 4:
            _write_vlc(w, *CHROMA_DC_TOTAL_ZEROS[total - 1], total_zeros)
        else:
            _write_vlc(w, *TOTAL_ZEROS[total - 1], total_zeros)
    zeros_left = total_zeros
    positions = [i for i, _ in reversed(nz)]
    for a, b in zip(positions, positions[1:]):
        if zeros_left == 0:
            break
        run = a - b - 1
        _write_vlc(w, *RUN_BEFORE[min(zeros_left, 7) - 1], run)
        zeros_left -= run
    return total


# ---------------------------------------------------------------- macroblock layer
def blk_pos(k):
    """(x, y) in 4x4 units of luma block index k (6.4.3)."""
    return (k // 4 % 2) * 2 + k % 2, (k // 8) * 2 + (k // 2) % 2


def blk_idx(x, y):
    return (y // 2) * 8 + (x // 2) * 4 + (y % 2) * 2 + x % 2


class Picture:
    """Encoder-side state of one picture: per macroblock the context the syntax of later
    macroblocks depends on (slice, nonzero counts, intra modes)."""

    def __init__(self, prm):
        self.prm = prm
        n = prm.width_mbs * prm.height_mbs
        self.slice = [-1] * n
        self.tc = np.zeros((n, 16), int)
        self.tcc = np.zeros((n, 2, 4), int)
        self.mode = np.full((n, 16), 2, int)

    def nb(self, mb, x, y):
        """(macroblock, 4x4 block) of the neighbour at (x, y) in 4x4 units, or None."""
        if 0 <= x < 4 and 0 <= y < 4:
            return mb, blk_idx(x, y)
        if x > 3 and y >= 0:
            return None
        W = self.prm.width_mbs
        mx, my = mb % W + (-1 if x < 0 else 1 if x > 3 else 0), mb // W + (-1 if y < 0 else 0)
        if mx < 0 or mx >= W or my < 0:
            return None
        a = my * W + mx
        if a >= mb or self.slice[a] != self.slice[mb]:
            return None
        return a, blk_idx(x % 4, y % 4)

    def luma_nc(self, mb, k):
        x, y = blk_pos(k)
        A, B = self.nb(mb, x - 1, y), self.nb(mb, x, y - 1)
        na = self.tc[A] if A else 0
        nb = self.tc[B] if B else 0
        if A and B:
            return (na + nb + 1) >> 1
        return na if A else nb if B else 0

    def chroma_nc(self, mb, c, k):
        x, y = k % 2, k // 2
        if x > 0:
            A = self.tcc[mb, c, k - 1]
        else:
            n = self.nb(mb, -1, 0)
            A = self.tcc[n[0], c, y * 2 + 1] if n else None
        if y > 0:
            B = self.tcc[mb, c, k - 2]
        else:
            n = self.nb(mb, 0, -1)
            B = self.tcc[n[0], c, 2 + x] if n else None
        if A is not None and B is not None:
            return (int(A) + int(B) + 1) >> 1
        return int(A) if A is not None else int(B) if B is not None else 0

    def mode_pred(self, mb, k):
        x, y = blk_pos(k)
        A, B = self.nb(mb, x - 1, y), self.nb(mb, x, y - 1)
        if not A or not B:
            return 2
        return min(self.mode[A], self.mode[B])

    def residual(self, w, mb, levels, cbp_luma, cbp_chroma, i16=False, t8=False):
        """levels: dict with 'luma' [16][16] scan order (Intra16x16: [16][15] AC), 'dc' [16],
        'luma8' [4][64] (8x8 transform), 'cdc' [2][4], 'cac' [2][4][15]."""
        if i16:
            encode_block(w, self.luma_nc(mb, 0), levels["dc"])
        for b8 in range(4):
            for i4 in range(4):
                k = b8 * 4 + i4
                if not cbp_luma & (1 << b8):
                    self.tc[mb, k] = 0
                    continue
                if i16:
                    t = encode_block(w, self.luma_nc(mb, k), levels["luma"][k])
                elif t8:
                    t = encode_block(w, self.luma_nc(mb, k), [levels["luma8"][b8][4 * j + i4] for j in range(16)])
                else:
                    t = encode_block(w, self.luma_nc(mb, k), levels["luma"][k])
                self.tc[mb, k] = t
        for c in range(2):
            if cbp_chroma & 3:
                encode_block(w, -1, levels["cdc"][c])
        for c in range(2):
            for k in range(4):
                if not cbp_chroma & 2:
                    self.tcc[mb, c, k] = 0
                    continue
                self.tcc[mb, c, k] = encode_block(w, self.chroma_nc(mb, c, k), levels["cac"][c][k])

    def pcm(self, w, mb, samples, slice_id):
        self.slice[mb] = slice_id
        w.align_zero()
        for s in samples:
            w.u(int(s), 8)
        self.tc[mb] = 16
        self.tcc[mb] = 16
        self.mode[mb] = 2

    def skip(self, mb, slice_id):
        self.slice[mb] = slice_id
        self.tc[mb] = 0
        self.tcc[mb] = 0
        self.mode[mb] = 2
