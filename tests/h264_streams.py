"""H.264 CAVLC streams through the oracle's front end (oracle/h264_entropy.c) into the library's
records: decoding access units, building Units for the device and the oracle, and the streams the
tests use -- the reference's weighted bi-prediction scenes (crates/codecs/h264/src/decoder.rs
:16853-16950, written here with tests/h264_bitstream.py) and seeded random multi-macroblock
streams whose syntax is known, so the front end's output can be compared with what was coded."""
import ctypes as C

import numpy as np

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

import h264_bitstream as hb


class PictureInfo(C.Structure):
    _fields_ = [("width_mbs", C.c_uint32), ("height_mbs", C.c_uint32), ("slot", C.c_int32), ("poc", C.c_int32),
                ("is_reference", C.c_int32), ("slice_type", C.c_uint32), ("ref_count", C.c_uint32 * 2),
                ("ref_slot", (C.c_int32 * 16) * 2), ("ref_poc", (C.c_int32 * 16) * 2), ("coded_mbs", C.c_uint32),
                ("weighted_mode", C.c_uint32), ("deblock", C.c_uint32), ("chroma_qp_offset", C.c_int32 * 2),
                ("transform_bypass", C.c_uint32)]


class Decoded:
    def __init__(self, info, inputs, mvd):
        self.info, self.inputs, self.mvd = info, inputs, mvd


class FrontEnd:
    def __init__(self, lib):
        self.lib = lib
        self.h = C.c_void_p(lib.oracle_h264_decoder_new())

    def close(self):
        if self.h:
            self.lib.oracle_h264_decoder_free(self.h)
            self.h = None

    def decode(self, au):
        """One access unit (list of NAL units): Decoded, None for parameter sets only, or the status."""
        data = hb.length_prefixed(au)
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
        info = PictureInfo()
        st = self.lib.oracle_h264_decode_access_unit(self.h, buf, len(data), C.byref(info))
        if st == 0:
            return None
        if st != 1:
            return st
        n = info.width_mbs * info.height_mbs
        mbs = np.zeros(n, wl.MB_DTYPE)
        blocks = np.zeros((n, 16), wl.BLOCK_DTYPE)
        levels = np.zeros((max(info.coded_mbs, 1), abi.H264_LEVELS_PER_MB), np.int16)
        intra = np.zeros(n, wl.INTRA_DTYPE)
        weights = np.zeros((2, abi.H264_MAX_REFS, 3), wl.WEIGHT_DTYPE)
        bi = np.zeros((abi.H264_MAX_REFS, abi.H264_MAX_REFS, 3), wl.BI_WEIGHT_DTYPE)
        mvd = np.zeros((n, 2, 16, 2), np.int16)
        assert self.lib.oracle_h264_picture_records(self.h, mbs.ctypes.data, blocks.ctypes.data, levels.ctypes.data,
                                                    intra.ctypes.data, weights.ctypes.data, bi.ctypes.data,
                                                    mvd.ctypes.data) == 0
        inputs = {"mbs": mbs, "blocks": blocks, "levels": levels, "weights": weights, "bi_weights": bi}
        if (mbs["flags"] & abi.H264_MB_INTRA).any():
            inputs["intra"] = intra
        return Decoded(info, inputs, mvd)


def unit_of(dec):
    """A workloads.Unit of a decoded picture; its references are named by front-end slot
    (fill's p["ref{list}_{i}_{plane}"] entries come from h264_refs)."""
    info = dec.info
    W, H = info.width_mbs * 16, info.height_mbs * 16
    outputs = {"y": ((H, W), np.uint8, None), "cb": ((H // 2, W // 2), np.uint8, None), "cr": ((H // 2, W // 2), np.uint8, None)}
    counts = [info.ref_count[0], info.ref_count[1]]
    slots = [[info.ref_slot[lst][i] for i in range(counts[lst])] for lst in range(2)]
    pocs = [[info.ref_poc[lst][i] for i in range(counts[lst])] for lst in range(2)]

    def fill(p):
        s = abi.new(abi.H264Picture, width_mbs=info.width_mbs, height_mbs=info.height_mbs, weighted_mode=info.weighted_mode,
                    deblock=1, macroblocks=p["mbs"], blocks=p["blocks"], levels=p["levels"], weights=p["weights"],
                    bi_weights=p["bi_weights"], intra=p.get("intra"), transform_bypass=info.transform_bypass)
        s.planes[0], s.planes[1], s.planes[2] = p["y"], p["cb"], p["cr"]
        s.chroma_qp_offset[0], s.chroma_qp_offset[1] = info.chroma_qp_offset[0], info.chroma_qp_offset[1]
        for lst in range(2):
            s.ref_count[lst] = counts[lst]
            for i in range(counts[lst]):
                ref = s.refs[lst][i]
                ref.poc = pocs[lst][i]
                for k, n in enumerate(("y", "cb", "cr")):
                    ref.planes[k] = p[f"ref{lst}_{i}_{n}"]
        return s

    return wl.Unit("h264", dec.inputs, outputs, fill, {"slots": slots, "slot": info.slot, "poc": info.poc})


def h264_refs(unit, planes_by_slot, kind):
    """fill() entries of a unit's references from planes kept per slot (host arrays or device pointers)."""
    extra = {}
    for lst in range(2):
        for i, slot in enumerate(unit.meta["slots"][lst]):
            for n in ("y", "cb", "cr"):
                v = planes_by_slot[slot][n]
                extra[f"ref{lst}_{i}_{n}"] = v.ctypes.data if kind == "host" else v
    return extra


def reconstruct_on_oracle(oracle, units):
    """Every picture of a stream on the oracle in decoding order; planes by slot."""
    planes = {}
    for u in units:
        out = u.host_outputs()
        s = u.host_struct(out, h264_refs(u, planes, "host"))
        assert oracle.oracle_h264_reconstruct(C.byref(s)) == 0
        planes[u.meta["slot"]] = out
    return planes


def decode_stream(oracle, aus):
    fe = FrontEnd(oracle)
    try:
        out = []
        for au in aus:
            d = fe.decode(au)
            assert not isinstance(d, int), f"front end status {d}"
            if d is not None:
                out.append(d)
        return out
    finally:
        fe.close()


# ---------------------------------------------------------------- the reference's bi-prediction scenes
def weighted_bipred_stream(mode):
    """IDR I_PCM (42, 90, 160), P I_PCM (100, 110, 120) at POC 4, then a B_Bi_16x16 macroblock with
    zero motion: mode "explicit" (weighted_bipred_idc 1, the weights of explicit_weighted_bi_slice,
    POC 2) or "implicit" (idc 2, POC 1). Expected B planes: explicit [88, 105, 150], implicit
    [57, 95, 150] (decoder.rs:16853-16950)."""
    idc = 1 if mode == "explicit" else 2
    prm = hb.Params(weighted_bipred_idc=idc)
    head = [hb.sps(max_refs=2), hb.pps(weighted_bipred_idc=idc)]
    w = hb.BitWriter()
    hb.slice_header(w, prm, 2, 0, 0, idr=True, nal_ref_idc=3)
    w.ue(25)
    hb.Picture(prm).pcm(w, 0, [42] * 256 + [90] * 64 + [160] * 64, 0)
    idr = hb.nal(3, 5, w.rbsp())
    w = hb.BitWriter()
    hb.slice_header(w, prm, 0, 1, 4, nal_ref_idc=2)
    w.ue(0)      # mb_skip_run
    w.ue(30)     # I_PCM in a P slice
    hb.Picture(prm).pcm(w, 0, [100] * 256 + [110] * 64 + [120] * 64, 0)
    p = hb.nal(2, 1, w.rbsp())
    w = hb.BitWriter()
    weights = None
    if mode == "explicit":
        weights = {"luma_denom": 1, "chroma_denom": 1, "l0": [(1, 10, [(1, 4), (3, -2)])], "l1": [(3, -6, [(3, -4), (1, 2)])]}
    hb.slice_header(w, prm, 1, 2, 2 if mode == "explicit" else 1, nal_ref_idc=0, weights=weights)
    w.ue(0)      # mb_skip_run
    w.ue(3)      # B_Bi_16x16
    for _ in range(4):
        w.se(0)  # mvd_l0, mvd_l1
    w.ue(0)      # coded_block_pattern 0
    b = hb.nal(0, 1, w.rbsp())
    return [head + [idr], [p], [b]]


# ---------------------------------------------------------------- seeded random streams
def _median(a, b, c):
    return sorted((a, b, c))[1]


class Motion:
    """Encoder-side motion of one picture (ref index and vector per list per 4x4 block) and the
    motion vector predictor of ITU-T H.264 8.4.1.3 / P_Skip 8.4.1.1, restated from the standard."""

    def __init__(self, pic):
        self.pic = pic
        n = len(pic.slice)
        self.ref = np.full((n, 2, 16), -1, int)
        self.mv = np.zeros((n, 2, 16, 2), int)
        self.intra = np.zeros(n, bool)

    def neighbour(self, mb, x, y, done):
        if 0 <= x < 4 and 0 <= y < 4:
            k = hb.blk_idx(x, y)
            return (mb, k) if done[k] else None
        return self.pic.nb(mb, x, y)

    def motion(self, n, lst):
        if n is None or self.intra[n[0]]:
            return -1, (0, 0)
        r = self.ref[n[0], lst, n[1]]
        return (r, tuple(self.mv[n[0], lst, n[1]])) if r >= 0 else (-1, (0, 0))

    def predict(self, mb, lst, ref, x, y, w, shape, done):
        A, B = self.neighbour(mb, x - 1, y, done), self.neighbour(mb, x, y - 1, done)
        Cn = self.neighbour(mb, x + w, y - 1, done)
        if Cn is None:
            Cn = self.neighbour(mb, x - 1, y - 1, done)
        (ra, ma), (rb, mbv), (rc, mc) = self.motion(A, lst), self.motion(B, lst), self.motion(Cn, lst)
        if B is None and Cn is None and A is not None:
            rb, mbv, rc, mc = ra, ma, ra, ma
        if shape == 1:
            if y == 0 and rb == ref:
                return mbv
            if y != 0 and ra == ref:
                return ma
        if shape == 2:
            if x == 0 and ra == ref:
                return ma
            if x != 0 and rc == ref:
                return mc
        hits = [m for r, m in ((ra, ma), (rb, mbv), (rc, mc)) if r == ref]
        if len(hits) == 1:
            return hits[0]
        return (_median(ma[0], mbv[0], mc[0]), _median(ma[1], mbv[1], mc[1]))

    def set(self, mb, lst, x, y, w, h, ref, mv, done=None):
        for yy in range(y, y + h):
            for xx in range(x, x + w):
                k = hb.blk_idx(xx, yy)
                self.ref[mb, lst, k] = ref
                self.mv[mb, lst, k] = mv

    def p_skip(self, mb):
        A, B = self.pic.nb(mb, -1, 0), self.pic.nb(mb, 0, -1)
        mv = (0, 0)
        if A is not None and B is not None:
            (ra, ma), (rb, mbv) = self.motion(A, 0), self.motion(B, 0)
            if not ((ra == 0 and ma == (0, 0)) or (rb == 0 and mbv == (0, 0))):
                mv = self.predict(mb, 0, 0, 0, 0, 4, 0, [0] * 16)
        self.set(mb, 0, 0, 0, 4, 4, 0, mv)
        return mv


def _avail(pic, mb):
    """(top, left, top-left) macroblock availability."""
    return pic.nb(mb, 0, -1) is not None, pic.nb(mb, -1, 0) is not None, pic.nb(mb, -1, -1) is not None


def _block_modes(rng, top, left, topleft):
    """Intra 4x4 / 8x8 modes usable with the neighbouring samples there are."""
    modes = [2]
    if top:
        modes += [0, 3, 7]
    if left:
        modes += [1, 8]
    if top and left and topleft:
        modes += [4, 5, 6]
    return int(rng.choice(modes))


def _sparse(rng, n, p=0.25, big=0.02):
    v = np.zeros(n, int)
    for i in range(n):
        if rng.random() < p:
            mag = int(rng.integers(1, 4)) if rng.random() > 0.2 else int(rng.integers(4, 40))
            if rng.random() < big:
                mag = int(rng.integers(100, 3000))
            v[i] = mag if rng.random() < 0.5 else -mag
    return v


def _levels(rng, cbp_luma, cbp_chroma, i16=False, t8=False):
    lv = {"luma": [_sparse(rng, 15 if i16 else 16) if cbp_luma & (1 << (k // 4)) else np.zeros(15 if i16 else 16, int)
                   for k in range(16)],
          "luma8": [_sparse(rng, 64, 0.15) if cbp_luma & (1 << b) else np.zeros(64, int) for b in range(4)],
          "dc": _sparse(rng, 16, 0.4),
          "cdc": [_sparse(rng, 4, 0.5) if cbp_chroma & 3 else np.zeros(4, int) for _ in range(2)],
          "cac": [[_sparse(rng, 15, 0.2) if cbp_chroma & 2 else np.zeros(15, int) for _ in range(4)] for _ in range(2)]}
    rec = np.zeros(384, int)
    if i16:
        for k in range(16):
            rec[16 * k] = lv["dc"][k]
            rec[16 * k + 1:16 * k + 16] = lv["luma"][k]
    elif t8:
        for b in range(4):
            rec[64 * b:64 * b + 64] = lv["luma8"][b]
    else:
        for k in range(16):
            rec[16 * k:16 * k + 16] = lv["luma"][k]
    for c in range(2):
        rec[256 + 4 * c:260 + 4 * c] = lv["cdc"][c]
        for k in range(4):
            rec[264 + 15 * (4 * c + k):279 + 15 * (4 * c + k)] = lv["cac"][c][k]
    return lv, rec


B_PRED = {1: (0,), 2: (1,), 3: (2,), 4: (0, 0), 5: (0, 0), 6: (1, 1), 7: (1, 1), 8: (0, 1), 9: (0, 1), 10: (1, 0),
          11: (1, 0), 12: (0, 2), 13: (0, 2), 14: (1, 2), 15: (1, 2), 16: (2, 0), 17: (2, 0), 18: (2, 1), 19: (2, 1),
          20: (2, 2), 21: (2, 2)}
B_SUB = {1: (0, 0), 2: (0, 1), 3: (0, 2), 4: (1, 0), 5: (2, 0), 6: (1, 1), 7: (2, 1), 8: (1, 2), 9: (2, 2), 10: (3, 0),
         11: (3, 1), 12: (3, 2)}   # sub_mb_type -> (shape 0 8x8 / 1 8x4 / 2 4x8 / 3 4x4, prediction)


def _write_intra(w, rng, pic, mb, slice_type, prm, exp):
    top, left, tl = _avail(pic, mb)
    kind = rng.choice(["i4", "i8", "i16", "pcm"], p=[0.35, 0.25 if prm.transform_8x8 else 0.0, 0.3,
                                                      0.1 + (0.25 if not prm.transform_8x8 else 0.0)])
    base = {2: 0, 0: 5, 1: 23}[slice_type]
    e = {"intra": True, "kind": kind}
    if kind == "pcm":
        w.ue(base + 25)
        samples = rng.integers(0, 256, 384)
        pic.pcm(w, mb, samples, pic.slice[mb])
        e.update(type=abi.H264_INTRA_PCM, cbp=0, record=samples, qp=0)
        exp[mb] = e
        return
    chroma_modes = [0] + ([1] if left else []) + ([2] if top else []) + ([3] if top and left and tl else [])
    chroma_mode = int(rng.choice(chroma_modes))
    if kind == "i16":
        modes = [2] + ([0] if top else []) + ([1] if left else []) + ([3] if top and left and tl else [])
        mode = int(rng.choice(modes))
        cbp_chroma = int(rng.integers(0, 3))
        cbp_luma = 15 if rng.random() < 0.5 else 0
        w.ue(base + 1 + mode + 4 * cbp_chroma + (12 if cbp_luma else 0))
        w.ue(chroma_mode)
        dq = int(rng.integers(-3, 4))
        w.se(dq)
        lv, rec = _levels(rng, cbp_luma, cbp_chroma, i16=True)
        pic.mode[mb] = 2
        pic.residual(w, mb, lv, cbp_luma, cbp_chroma, i16=True)
        e.update(type=abi.H264_INTRA_16X16, luma_mode=mode, chroma_mode=chroma_mode, cbp=cbp_luma | (cbp_chroma << 4),
                 record=rec, dq=dq)
        exp[mb] = e
        return
    t8 = kind == "i8"
    w.ue(base)
    if prm.transform_8x8:
        w.bit(t8)
    modes = []
    for i in range(4 if t8 else 16):
        k = 4 * i if t8 else i
        x, y = hb.blk_pos(k)
        btop = y > 0 or top
        bleft = x > 0 or left
        btl = (x > 0 and y > 0) or (x > 0 and top) or (y > 0 and left) or tl
        mode = _block_modes(rng, btop, bleft, btl)
        pred = pic.mode_pred(mb, k)
        if mode == pred:
            w.bit(1)
        else:
            w.bit(0)
            w.u(mode if mode < pred else mode - 1, 3)
        for j in range(4 if t8 else 1):
            pic.mode[mb, k + j] = mode
        modes.append(mode)
    w.ue(chroma_mode)
    cbp = int(rng.choice(hb.CBP_INTRA))
    w.ue(hb.CBP_INTRA.index(cbp))
    dq = 0
    if cbp:
        dq = int(rng.integers(-3, 4))
        w.se(dq)
    lv, rec = _levels(rng, cbp & 15, cbp >> 4, t8=t8)
    pic.residual(w, mb, lv, cbp & 15, cbp >> 4, t8=t8)
    e.update(type=abi.H264_INTRA_8X8 if t8 else abi.H264_INTRA_4X4, chroma_mode=chroma_mode, modes=modes, cbp=cbp,
             record=rec if cbp else None, dq=dq, t8=t8)
    exp[mb] = e


def _write_inter(w, rng, pic, mot, mb, slice_type, prm, active, exp):
    done = [0] * 16
    e = {"intra": False, "ref": np.full((2, 16), -1, int), "mvd": np.zeros((2, 16, 2), int), "mv": None}
    if slice_type == 0:
        t = int(rng.integers(0, 5))
        w.ue(t)
        sub = t >= 3
        shape = t if t < 3 else None
        preds = (0,) if t == 0 else (0, 0)
    else:
        t = int(rng.integers(1, 23))
        w.ue(t)
        sub = t == 22
        shape = None if sub else (0 if t <= 3 else 1 if t % 2 == 0 else 2)
        preds = None if sub else B_PRED[t]
    allow8 = True
    if not sub:
        parts = len(preds)
        refs = [[0, 0] for _ in range(2)]
        for lst in range(2):
            for i in range(parts):
                if preds[i] in (lst, 2) and active[lst] > 1:
                    refs[lst][i] = int(rng.integers(0, active[lst]))
                    w.te(refs[lst][i], active[lst] - 1)
        mvds = [[None, None] for _ in range(2)]
        for lst in range(2):
            for i in range(parts):
                if preds[i] in (lst, 2):
                    mvds[lst][i] = (int(rng.integers(-12, 13)), int(rng.integers(-12, 13)))
                    w.se(mvds[lst][i][0])
                    w.se(mvds[lst][i][1])
        for i in range(parts):
            x, y = (2 * i if shape == 2 else 0), (2 * i if shape == 1 else 0)
            bw, bh = (2 if shape == 2 else 4), (2 if shape == 1 else 4)
            for lst in range(2):
                if preds[i] not in (lst, 2):
                    continue
                p = mot.predict(mb, lst, refs[lst][i], x, y, bw, shape, done)
                mv = (p[0] + mvds[lst][i][0], p[1] + mvds[lst][i][1])
                mot.set(mb, lst, x, y, bw, bh, refs[lst][i], mv)
                for yy in range(y, y + bh):
                    for xx in range(x, x + bw):
                        k = hb.blk_idx(xx, yy)
                        e["ref"][lst, k] = refs[lst][i]
                        e["mvd"][lst, k] = mvds[lst][i]
            for yy in range(y, y + bh):
                for xx in range(x, x + bw):
                    done[hb.blk_idx(xx, yy)] = 1
    else:
        subs = [int(rng.integers(0, 4)) if slice_type == 0 else int(rng.integers(1, 13)) for _ in range(4)]
        for s in subs:
            w.ue(s)
        shp = [s if slice_type == 0 else B_SUB[s][0] for s in subs]
        prd = [0 if slice_type == 0 else B_SUB[s][1] for s in subs]
        refs = [[0] * 4 for _ in range(2)]
        for lst in range(2):
            for i in range(4):
                if prd[i] in (lst, 2) and active[lst] > 1 and not (slice_type == 0 and t == 4):
                    refs[lst][i] = int(rng.integers(0, active[lst]))
                    w.te(refs[lst][i], active[lst] - 1)
        mvds = {}
        for lst in range(2):
            for i in range(4):
                if prd[i] not in (lst, 2):
                    continue
                for j in range(1 if shp[i] == 0 else 4 if shp[i] == 3 else 2):
                    mvds[(lst, i, j)] = (int(rng.integers(-12, 13)), int(rng.integers(-12, 13)))
                    w.se(mvds[(lst, i, j)][0])
                    w.se(mvds[(lst, i, j)][1])
        for i in range(4):
            x8, y8 = (i % 2) * 2, (i // 2) * 2
            n = 1 if shp[i] == 0 else 4 if shp[i] == 3 else 2
            bw, bh = (2 if shp[i] in (0, 1) else 1), (2 if shp[i] in (0, 2) else 1)
            for j in range(n):
                x = x8 + (j % 2 if shp[i] in (2, 3) else 0)
                y = y8 + (j if shp[i] == 1 else j // 2 if shp[i] == 3 else 0)
                for lst in range(2):
                    if prd[i] not in (lst, 2):
                        continue
                    p = mot.predict(mb, lst, refs[lst][i], x, y, bw, 0, done)
                    mv = (p[0] + mvds[(lst, i, j)][0], p[1] + mvds[(lst, i, j)][1])
                    mot.set(mb, lst, x, y, bw, bh, refs[lst][i], mv)
                    for yy in range(y, y + bh):
                        for xx in range(x, x + bw):
                            k = hb.blk_idx(xx, yy)
                            e["ref"][lst, k] = refs[lst][i]
                            e["mvd"][lst, k] = mvds[(lst, i, j)]
                for yy in range(y, y + bh):
                    for xx in range(x, x + bw):
                        done[hb.blk_idx(xx, yy)] = 1
        allow8 = all(s == 0 for s in shp)
    cbp = int(rng.choice(hb.CBP_INTER))
    w.ue(hb.CBP_INTER.index(cbp))
    t8 = False
    if (cbp & 15) and prm.transform_8x8 and allow8:
        t8 = bool(rng.random() < 0.5)
        w.bit(t8)
    dq = 0
    if cbp:
        dq = int(rng.integers(-3, 4))
        w.se(dq)
        lv, rec = _levels(rng, cbp & 15, cbp >> 4, t8=t8)
        pic.residual(w, mb, lv, cbp & 15, cbp >> 4, t8=t8)
    else:
        rec = None
        pic.tc[mb] = 0
        pic.tcc[mb] = 0
    pic.mode[mb] = 2
    e.update(cbp=cbp, record=rec, dq=dq, t8=t8, mv=mot.mv[mb].copy())
    exp[mb] = e


def random_stream(seed, width_mbs=5, height_mbs=4, pattern="IPBPB", high=True, weighted=True, max_slices=3):
    """A seeded stream of CAVLC pictures of every supported macroblock kind. Returns the access units
    and, per picture, the coded syntax per macroblock (kind, cbp, QP delta, level record, intra modes,
    reference indices, mvds, the encoder's motion vectors)."""
    rng = np.random.default_rng(seed)
    prm = hb.Params(width_mbs, height_mbs, log2_max_frame_num=4, log2_max_poc_lsb=6,
                    weighted_pred=weighted and rng.random() < 0.5,
                    weighted_bipred_idc=int(rng.integers(0, 3)) if weighted else 0, init_qp=int(rng.integers(20, 36)),
                    transform_8x8=high, num_ref_idx_default=(2, 2))
    head = [hb.sps(width_mbs, height_mbs, max_refs=3, profile=100 if high else 77, log2_max_poc_lsb=6),
            hb.pps(weighted_pred=prm.weighted_pred, weighted_bipred_idc=prm.weighted_bipred_idc, init_qp=prm.init_qp,
                   transform_8x8=high, num_ref_idx_default=(2, 2), chroma_qp_offset=int(rng.integers(-2, 3)))]
    n = width_mbs * height_mbs
    aus, expected = [], []
    frame_num, refs_held, poc = 0, 0, 0
    for pi, ptype in enumerate(pattern):
        slice_type = {"I": 2, "P": 0, "B": 1}[ptype]
        idr = pi == 0
        is_ref = ptype != "B"
        poc = 0 if idr else (poc + 8 if ptype == "P" else poc)
        this_poc = poc - 4 if ptype == "B" else poc
        pic = hb.Picture(prm)
        mot = Motion(pic)
        exp = [None] * n
        cuts = sorted(rng.choice(np.arange(1, n), int(rng.integers(0, max_slices)), replace=False).tolist()) if n > 1 else []
        starts = [0] + cuts
        nals = []
        pic_weights = None
        if (prm.weighted_pred and slice_type == 0) or (prm.weighted_bipred_idc == 1 and slice_type == 1):
            ld, cd = int(rng.integers(0, 7)), int(rng.integers(0, 7))

            def entry():
                if rng.random() < 0.3:
                    return None
                return (int(rng.integers(-20, 90)), int(rng.integers(-30, 31)),
                        None if rng.random() < 0.3 else [(int(rng.integers(-20, 90)), int(rng.integers(-30, 31)))
                                                         for _ in range(2)])
            pic_weights = {"luma_denom": ld, "chroma_denom": cd, "l0": [entry() for _ in range(3)],
                           "l1": [entry() for _ in range(3)]}
        for si, first in enumerate(starts):
            last = (starts[si + 1] if si + 1 < len(starts) else n)
            w = hb.BitWriter()
            qp = int(np.clip(prm.init_qp + rng.integers(-4, 5), 0, 51))
            num_ref = None
            if slice_type != 2:
                avail = max(1, min(refs_held, 3))
                num_ref = (int(rng.integers(1, avail + 1)), int(rng.integers(1, avail + 1)) if slice_type == 1 else 1)
            weights = pic_weights   # one table per picture (the records carry one)
            dis = int(rng.integers(0, 3))
            active = hb.slice_header(w, prm, slice_type, frame_num, this_poc % 64, idr=idr, nal_ref_idc=2 if is_ref else 0,
                                     first_mb=first, num_ref_idx=num_ref, weights=weights, qp=qp, disable_deblocking=dis,
                                     alpha=int(rng.integers(-3, 4)) * 2, beta=int(rng.integers(-3, 4)) * 2)
            running_qp = qp
            mb = first
            while mb < last:
                pic.slice[mb] = si
                if slice_type == 0 and rng.random() < 0.2:
                    run = int(min(rng.integers(1, 4), last - mb))
                    w.ue(run)
                    for _ in range(run):
                        pic.skip(mb, si)
                        mv = mot.p_skip(mb)
                        exp[mb] = {"intra": False, "skip": True, "cbp": 0, "qp": running_qp, "slice": si,
                                   "mv": mot.mv[mb].copy(), "ref": mot.ref[mb].copy(), "mvd": np.zeros((2, 16, 2), int)}
                        mb += 1
                    if mb >= last:
                        break
                    pic.slice[mb] = si
                elif slice_type != 2:
                    w.ue(0)
                if slice_type == 2 or rng.random() < 0.25:
                    mot.intra[mb] = True
                    _write_intra(w, rng, pic, mb, slice_type, prm, exp)
                else:
                    _write_inter(w, rng, pic, mot, mb, slice_type, prm, active, exp)
                e = exp[mb]
                e["slice"] = si
                if e.get("kind") == "pcm":
                    e["qp"] = 0
                else:
                    running_qp = (running_qp + e.get("dq", 0) + 52) % 52
                    e["qp"] = running_qp
                mb += 1
            nals.append(hb.nal(2 if is_ref else 0, 5 if idr else 1, w.rbsp()))
        aus.append((head if pi == 0 else []) + nals)
        expected.append(exp)
        if is_ref:
            frame_num = (frame_num + 1) % 16
            refs_held = min(refs_held + 1, 3)
    return aus, expected
