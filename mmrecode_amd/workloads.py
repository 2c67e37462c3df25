"""Seeded synthetic post-entropy workloads for the reconstruction path.

Every frame / picture is a `Unit`: named numpy input arrays, named output shapes, and a function
that fills the include/mmr_recon.h descriptor from a name -> pointer map. The same Unit therefore
drives the oracle (host pointers) and the device (device pointers) on identical inputs.

Generators follow the workload definitions of the benchmark configurations:
  mjpeg_frame   640x480 4:2:0 baseline JPEG coefficient blocks
  mpeg2_gop     1920x1080 closed GOP of 15 (decode order I P B B P B B P B B P B B B B)
  dv_frame      DV25 625/50 (4:2:0) or 525/60 (4:1:1) frames
  h264_picture  inter pictures with 4 resident references per resolution
Randomness comes from numpy's PCG64 seeded per unit, so a (seed, index) pair always yields the
same bytes.
"""
import ctypes as C

import numpy as np

from . import abi


class Unit:
    def __init__(self, kind, inputs, outputs, fill, meta=None):
        self.kind = kind
        self.inputs = inputs          # name -> np.ndarray
        self.outputs = outputs        # name -> (shape, dtype, init value or None)
        self.fill = fill              # fill(ptrs) -> abi struct
        self.meta = meta or {}

    def input_bytes(self):
        return sum(a.nbytes for a in self.inputs.values())

    def output_bytes(self):
        return sum(int(np.prod(s)) * np.dtype(d).itemsize for s, d, _ in self.outputs.values())

    # ---- host (oracle) side
    def host_outputs(self):
        out = {}
        for name, (shape, dtype, init) in self.outputs.items():
            out[name] = np.full(shape, init, dtype) if init is not None else np.zeros(shape, dtype)
        return out

    def host_struct(self, outputs, extra=None):
        ptrs = {k: v.ctypes.data for k, v in self.inputs.items()}
        ptrs.update({k: v.ctypes.data for k, v in outputs.items()})
        if extra:
            ptrs.update(extra)
        return self.fill(ptrs)


class DeviceUnit:
    """A Unit resident on the device: uploaded inputs + allocated outputs."""

    def __init__(self, ctx, unit, extra=None):
        self.ctx = ctx
        self.unit = unit
        self.bufs = {k: ctx.upload(v) for k, v in unit.inputs.items()}
        for name, (shape, dtype, init) in unit.outputs.items():
            nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
            buf = ctx.alloc(nbytes)
            if init is not None:
                ctx.write(buf, np.full(shape, init, dtype))
            self.bufs[name] = buf
        ptrs = {k: b.ptr for k, b in self.bufs.items()}
        if extra:
            ptrs.update(extra)
        self.struct = unit.fill(ptrs)

    def ptr(self, name):
        return self.bufs[name].ptr

    def download(self, name):
        shape, dtype, _ = self.unit.outputs[name]
        return self.ctx.download(self.bufs[name], dtype, shape)

    def free(self):
        for b in self.bufs.values():
            b.free()


def _round_half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def laplace_levels(rng, n, scale0=12.0, decay=10.0, zero_fraction=0.85, first=1, length=None):
    """Levels for scan indices k = first .. first+length-1 (default up to 63):
    round_half_away(Laplace(0, scale0 * exp(-k / decay))), thinned with one keep probability so
    that about `zero_fraction` of them are zero overall. Returns int16 [n, length]."""
    if length is None:
        length = 64 - first
    k = np.arange(first, first + length)
    scale = scale0 * np.exp(-k / decay)
    v = _round_half_away(rng.laplace(0.0, 1.0, size=(n, k.size)) * scale)
    p_nonzero = np.exp(-0.5 / scale).mean()
    keep = min(1.0, (1.0 - zero_fraction) / max(p_nonzero, 1e-9))
    v *= rng.random((n, k.size)) < keep
    return np.clip(v, -32768, 32767).astype(np.int16)


# ---------------------------------------------------------------- MJPEG (config 0)
def mjpeg_frame(seed, index=0, width=640, height=480, sampling=((2, 2), (1, 1), (1, 1)), rgb=True):
    rng = np.random.Generator(np.random.PCG64([seed, 0, index]))
    ncomp = len(sampling)
    hmax = max(h for h, _ in sampling)
    vmax = max(v for _, v in sampling)
    interleaved = ncomp > 1
    mcu_w, mcu_h = (hmax * 8, vmax * 8) if interleaved else (8, 8)
    cols, rows = -(-width // mcu_w), -(-height // mcu_h)
    per_mcu = sum(h * v for h, v in sampling) if interleaved else 1
    nblocks = cols * rows * per_mcu
    levels = np.zeros((nblocks, 64), np.int16)
    levels[:, 0] = rng.integers(-128, 128, nblocks)
    levels[:, 1:] = laplace_levels(rng, nblocks)
    quant = np.zeros((2, 64), np.uint16)
    for t in range(2):
        quant[t] = np.sort(rng.integers(2, 65, 64))
    comp_dims = [(-(-width * h // hmax), -(-height * v // vmax)) for h, v in sampling]
    outputs = {f"plane{c}": ((comp_dims[c][1], comp_dims[c][0]), np.uint8, None) for c in range(ncomp)}
    if rgb and ncomp == 3:
        outputs["rgb"] = ((height, width, 3), np.uint8, None)

    def fill(p):
        f = abi.new(abi.MjpegFrame, width=width, height=height, component_count=ncomp, block_count=nblocks,
                    levels=p["levels"], quant_tables=p["quant"], rgb=p.get("rgb"), rgb_range=abi.RANGE_FULL)
        for c, (h, v) in enumerate(sampling):
            f.h_sampling[c] = h
            f.v_sampling[c] = v
            f.quant_table[c] = 0 if c == 0 else 1
            f.planes[c] = p[f"plane{c}"]
        return f

    return Unit("mjpeg", {"levels": levels, "quant": quant}, outputs, fill,
                {"width": width, "height": height, "blocks": nblocks})


def mjpeg_unit_from_frame(frame, rgb=True):
    """A Unit holding one entropy-decoded JPEG frame (a dict of oracle.binding.jpeg_frames):
    its levels, quantisation tables and geometry, with outputs sized like the reference's planes."""
    width, height, sampling = frame["width"], frame["height"], frame["sampling"]
    ncomp = len(sampling)
    hmax = max(h for h, _ in sampling)
    vmax = max(v for _, v in sampling)
    comp_dims = [(-(-width * h // hmax), -(-height * v // vmax)) for h, v in sampling]
    outputs = {f"plane{c}": ((comp_dims[c][1], comp_dims[c][0]), np.uint8, None) for c in range(ncomp)}
    if rgb and ncomp == 3:
        outputs["rgb"] = ((height, width, 3), np.uint8, None)
    levels = np.ascontiguousarray(frame["levels"], np.int16)

    def fill(p):
        f = abi.new(abi.MjpegFrame, width=width, height=height, component_count=ncomp, block_count=len(levels),
                    levels=p["levels"], quant_tables=p["quant"], rgb=p.get("rgb"), rgb_range=abi.RANGE_FULL)
        for c, (h, v) in enumerate(sampling):
            f.h_sampling[c] = h
            f.v_sampling[c] = v
            f.quant_table[c] = frame["qsel"][c]
            f.planes[c] = p[f"plane{c}"]
        return f

    return Unit("mjpeg", {"levels": levels, "quant": np.ascontiguousarray(frame["quant"], np.uint16)}, outputs, fill,
                {"width": width, "height": height, "blocks": len(levels)})


# ---------------------------------------------------------------- MPEG-2 (config 1)
GOP_DECODE_ORDER = "IPBBPBBPBBPBBBB"   # display I B B P B B P B B P B B P B B, closed


# mmr_recon_mpeg2_macroblock as a numpy record
MPEG2_MB_DTYPE = np.dtype([("flags", "u1"), ("cbp", "u1"), ("qscale", "u1"), ("reserved", "u1"),
                           ("mv", "<i2", (2, 2)), ("block_offset", "<u4")])
# mmr_recon_mpeg2_field_motion
MPEG2_FIELD_DTYPE = np.dtype([("mv", "<i2", (2, 2)), ("field_select", "u1", (2, 2))])


def mpeg2_unit(width, height, picture_type, matrices, mbs, levels, field_motion=None, alternate_scan=0,
               intra_dc_precision=0, fwd=None, bwd=None, meta=None):
    """One MPEG-2 frame picture as a Unit: picture_type is MMR_MPEG2_PICTURE_*, mbs the raster-order
    macroblock records (MPEG2_MB_DTYPE), levels int16 [blocks][64], field_motion MPEG2_FIELD_DTYPE
    records or None; fwd / bwd are the decode-order indices of its references (unit.meta)."""
    mb_w, mb_h = -(-width // 16), -(-height // 16)
    W, H = mb_w * 16, mb_h * 16
    inputs = {"matrices": np.ascontiguousarray(matrices, np.uint8), "mbs": np.ascontiguousarray(mbs),
              "levels": np.ascontiguousarray(levels, np.int16)}
    if field_motion is not None:
        inputs["field_motion"] = np.ascontiguousarray(field_motion)
    outputs = {"y": ((H, W), np.uint8, None), "cb": ((H // 2, W // 2), np.uint8, None),
               "cr": ((H // 2, W // 2), np.uint8, None)}

    def fill(p):
        s = abi.new(abi.Mpeg2Picture, width=width, height=height, picture_type=picture_type,
                    alternate_scan=alternate_scan, intra_dc_precision=intra_dc_precision, matrices=p["matrices"],
                    macroblocks=p["mbs"], levels=p["levels"], field_motion=p.get("field_motion"),
                    forward_source=p.get("forward_source", -1), backward_source=p.get("backward_source", -1))
        for k, n in enumerate(("y", "cb", "cr")):
            s.planes[k] = p[n]
            s.forward[k] = p.get("fwd_" + n)
            s.backward[k] = p.get("bwd_" + n)
        return s

    m = {"fwd": fwd, "bwd": bwd, "width": width, "height": height}
    m.update(meta or {})
    return Unit("mpeg2", inputs, outputs, fill, m)


def mpeg2_gop(seed, gop_index=0, width=1920, height=1080, gop=GOP_DECODE_ORDER, alt_scan_fraction=0.3,
              field_fraction=0.0):
    """One closed GOP in decode order. Returns a list of Units; unit.meta has 'type', 'fwd', 'bwd'
    (indices of reference pictures inside this GOP, or None). field_fraction of the predicted
    macroblocks use field prediction (random field vectors, some reaching past the picture edges,
    and random field selects)."""
    mb_w, mb_h = -(-width // 16), -(-height // 16)
    n_mb = mb_w * mb_h
    W, H = mb_w * 16, mb_h * 16
    units = []
    older = newer = None
    for pi, ptype in enumerate(gop):
        rng = np.random.Generator(np.random.PCG64([seed, 1, gop_index, pi]))
        matrices = rng.integers(8, 84, (4, 64)).astype(np.uint8)
        alt = int(rng.random() < alt_scan_fraction)
        u = rng.random(n_mb)
        if ptype == "I":
            kind = np.zeros(n_mb, np.int64)                      # 0 intra
        elif ptype == "P":
            kind = np.where(u < 0.15, 0, 1)                      # intra / forward
        else:
            kind = np.select([u < 0.15, u < 0.50, u < 0.65], [0, 1, 2], 3)  # intra / fwd / bwd / bi
        flags = np.zeros(n_mb, np.uint8)
        flags[kind == 0] = abi.MPEG2_MB_INTRA
        flags[kind == 1] = abi.MPEG2_MB_FORWARD
        flags[kind == 2] = abi.MPEG2_MB_BACKWARD
        flags[kind == 3] = abi.MPEG2_MB_FORWARD | abi.MPEG2_MB_BACKWARD
        cbp = np.where(kind == 0, 0x3F, rng.integers(0, 64, n_mb)).astype(np.uint8)
        qscale = (2 * rng.integers(2, 32, n_mb)).astype(np.uint8)
        mbx = np.arange(n_mb) % mb_w
        mby = np.arange(n_mb) // mb_w
        mv = rng.integers(-128, 129, (n_mb, 2, 2))
        # keep the 16x16 prediction inside the coded picture
        lo_x, hi_x = -2 * mbx * 16, 2 * (W - 16 - mbx * 16)
        lo_y, hi_y = -2 * mby * 16, 2 * (H - 16 - mby * 16)
        mv[:, :, 0] = np.clip(mv[:, :, 0], lo_x[:, None], hi_x[:, None])
        mv[:, :, 1] = np.clip(mv[:, :, 1], lo_y[:, None], hi_y[:, None])
        field_motion = None
        if field_fraction > 0 and ptype != "I":
            field = (kind != 0) & (rng.random(n_mb) < field_fraction)
            flags[field] |= abi.MPEG2_MB_FIELD_MOTION
            field_motion = np.zeros(n_mb, dtype=MPEG2_FIELD_DTYPE)
            fv = rng.integers(-48, 49, (n_mb, 2, 2, 2))          # [mb][parity][direction][x, y]
            mv[field] = fv[field, 0]
            field_motion["mv"] = fv[:, 1]
            field_motion["field_select"] = rng.integers(0, 2, (n_mb, 2, 2))
            interlaced = (kind != 0) & (rng.random(n_mb) < 0.5)   # field DCT on half of them
            flags[interlaced] |= abi.MPEG2_MB_INTERLACED_DCT
        coded_per_mb = np.array([bin(int(c)).count("1") for c in range(64)])[cbp & 0x3F]
        offsets = np.concatenate([[0], np.cumsum(coded_per_mb)[:-1]]).astype(np.uint32)
        ncoded = int(coded_per_mb.sum())
        levels = np.zeros((max(ncoded, 1), 64), np.int16)
        if ncoded:
            levels[:ncoded, 1:] = laplace_levels(rng, ncoded)
            levels[:ncoded, 0] = _round_half_away(rng.laplace(0, 12.0, ncoded)).astype(np.int16)
            # intra blocks: DC predictor value (intra_dc_precision 0: 8-bit DC)
            intra_blocks = np.repeat(kind == 0, coded_per_mb)
            levels[:ncoded][intra_blocks, 0] = rng.integers(0, 256, int(intra_blocks.sum()))
        mbs = np.zeros(n_mb, dtype=MPEG2_MB_DTYPE)
        mbs["flags"] = flags
        mbs["cbp"] = cbp
        mbs["qscale"] = qscale
        mbs["mv"] = mv
        mbs["block_offset"] = offsets
        ptype_code = {"I": abi.MPEG2_PICTURE_I, "P": abi.MPEG2_PICTURE_P, "B": abi.MPEG2_PICTURE_B}[ptype]
        fwd = bwd = None
        if ptype == "P":
            fwd = newer
        elif ptype == "B":
            fwd, bwd = older, newer
        units.append(mpeg2_unit(width, height, ptype_code, matrices, mbs, levels, field_motion, alt, 0, fwd, bwd,
                                {"type": ptype, "coded_blocks": ncoded, "kinds": np.bincount(kind, minlength=4)}))
        if ptype in "IP":
            older, newer = newer, pi
    return units


def mpeg2_ref_ptrs(outputs_by_index, unit, kind):
    """name -> pointer entries for a picture's references, from the output pointer maps of the
    pictures it references (kind 'host' arrays or device pointer dicts)."""
    extra = {}
    for tag, ref in (("fwd", unit.meta["fwd"]), ("bwd", unit.meta["bwd"])):
        if ref is None:
            continue
        src = outputs_by_index[ref]
        for n in ("y", "cb", "cr"):
            v = src[n]
            extra[f"{tag}_{n}"] = v.ctypes.data if kind == "host" else v
    return extra


# ---------------------------------------------------------------- DV (config 2)
def dv_frame(seed, index=0, system=abi.DV_SYSTEM_625_50, mode248_fraction=0.3, zero_fraction=0.8):
    rng = np.random.Generator(np.random.PCG64([seed, 2, index]))
    seqs = 12 if system == abi.DV_SYSTEM_625_50 else 10
    nblocks = seqs * 27 * 5 * 6
    height = 576 if system == abi.DV_SYSTEM_625_50 else 480
    cw, ch = (360, 288) if system == abi.DV_SYSTEM_625_50 else (180, 480)
    levels = np.zeros((nblocks, 64), np.int16)
    levels[:, 0] = rng.integers(-256, 256, nblocks)
    mag = rng.geometric(1.0 / 3.0, (nblocks, 63))
    sign = np.where(rng.random((nblocks, 63)) < 0.5, -1, 1)
    nz = rng.random((nblocks, 63)) >= zero_fraction
    levels[:, 1:] = (mag * sign * nz).astype(np.int16)
    mode = (rng.random(nblocks) < mode248_fraction).astype(np.uint8)
    cls = rng.integers(0, 4, nblocks).astype(np.uint8)
    qno = rng.integers(0, 16, nblocks).astype(np.uint8)
    info = (mode | (cls << 1) | (qno << 3)).astype(np.uint8)
    return dv_unit(system, levels, info)


def dv_unit(system, levels, info):
    """One DV25 frame as a Unit: levels int16 [blocks][64] by scan position ([0] = DC) and info
    uint8 [blocks] in DIF order (sequence, slot, macroblock, block)."""
    seqs = 12 if system == abi.DV_SYSTEM_625_50 else 10
    nblocks = seqs * 27 * 5 * 6
    assert levels.shape == (nblocks, 64) and info.shape == (nblocks,)
    height = 576 if system == abi.DV_SYSTEM_625_50 else 480
    cw, ch = (360, 288) if system == abi.DV_SYSTEM_625_50 else (180, 480)
    outputs = {"y": ((height, 720), np.uint8, None), "cb": ((ch, cw), np.uint8, None), "cr": ((ch, cw), np.uint8, None)}

    def fill(p):
        s = abi.new(abi.DvFrame, system=system, levels=p["levels"], block_info=p["info"])
        s.planes[0], s.planes[1], s.planes[2] = p["y"], p["cb"], p["cr"]
        return s

    return Unit("dv", {"levels": np.ascontiguousarray(levels, np.int16), "info": np.ascontiguousarray(info, np.uint8)},
                outputs, fill, {"system": system, "blocks": nblocks, "width": 720, "height": height})


# ---------------------------------------------------------------- H.264 (config 3)
MB_DTYPE = np.dtype([("flags", "u1"), ("qp", "u1"), ("cbp", "u1"), ("reserved", "u1"), ("slice_id", "<u2"),
                     ("offset_a", "i1"), ("offset_b", "i1"), ("coef_offset", "<u4")])
BLOCK_DTYPE = np.dtype([("mv", "<i2", (2, 2)), ("ref_idx", "i1", (2,)), ("total_coeff", "u1"), ("reserved", "u1")])
WEIGHT_DTYPE = np.dtype([("weight", "<i2"), ("offset", "<i2"), ("log2_denom", "u1"), ("reserved", "u1", (3,))])
BI_WEIGHT_DTYPE = np.dtype([("w0", "<i2"), ("w1", "<i2"), ("o0", "<i2"), ("o1", "<i2"), ("log2_denom", "u1"),
                            ("reserved", "u1", (3,))])

# 4x4 block index -> (x, y) in 4x4-block units (luma_block_position)
BLOCK_XY = np.array([[((k >> 2) & 1) * 2 + (k & 1), (k >> 3) * 2 + ((k >> 1) & 1)] for k in range(16)])


def h264_reference_planes(seed, width_mbs, height_mbs, index):
    """Smooth synthetic reference picture (gradients + low-amplitude noise) so that deblocking
    thresholds are met on a useful share of edges."""
    rng = np.random.Generator(np.random.PCG64([seed, 30, width_mbs, height_mbs, index]))
    W, H = width_mbs * 16, height_mbs * 16

    def plane(w, h):
        yy, xx = np.mgrid[0:h, 0:w]
        fx, fy = rng.uniform(0.002, 0.02, 2)
        base = 128 + 60 * np.sin(xx * fx + rng.uniform(0, 6)) * np.cos(yy * fy + rng.uniform(0, 6))
        return np.clip(base + rng.normal(0, 3, (h, w)), 0, 255).astype(np.uint8)

    return plane(W, H), plane(W // 2, H // 2), plane(W // 2, H // 2)


# mmr_recon_h264_intra
INTRA_DTYPE = np.dtype([("type", "u1"), ("luma_mode", "u1"), ("chroma_mode", "u1"), ("reserved", "u1"),
                        ("modes", "u1", (16,))])

# Intra4x4 / Intra8x8 modes by the neighbours they need
_NXN_MODES = {(False, False): [2], (True, False): [0, 2, 3, 7], (False, True): [1, 2, 8],
              (True, True): list(range(9))}
_I16_MODES = {(False, False): [2], (True, False): [0, 2], (False, True): [1, 2], (True, True): [0, 1, 2, 3]}
_CHROMA_MODES = {(False, False): [0], (True, False): [0, 2], (False, True): [0, 1], (True, True): [0, 1, 2, 3]}


def _intra_records(rng, intra, slice_id, width_mbs, type_weights):
    """Intra records for the macroblocks flagged in `intra`, with modes whose neighbours exist
    (same slice, inside the picture); 4x4 / 8x8 blocks inside the macroblock have the neighbours
    the macroblock's own earlier blocks give."""
    n_mb = intra.size
    rec = np.zeros(n_mb, INTRA_DTYPE)
    itype = rng.choice(4, n_mb, p=type_weights)
    for a in np.nonzero(intra)[0]:
        x, y = a % width_mbs, a // width_mbs
        left = bool(x > 0 and slice_id[a - 1] == slice_id[a])
        top = bool(y > 0 and slice_id[a - width_mbs] == slice_id[a])
        t = int(itype[a])
        rec["type"][a] = t
        rec["chroma_mode"][a] = rng.choice(_CHROMA_MODES[(top, left)])
        if t == abi.H264_INTRA_16X16:
            rec["luma_mode"][a] = rng.choice(_I16_MODES[(top, left)])
        elif t == abi.H264_INTRA_4X4:
            for k in range(16):
                bx, by = BLOCK_XY[k]
                rec["modes"][a, k] = rng.choice(_NXN_MODES[(bool(by > 0 or top), bool(bx > 0 or left))])
        elif t == abi.H264_INTRA_8X8:
            for g in range(4):
                rec["modes"][a, g] = rng.choice(_NXN_MODES[(bool(g >= 2 or top), bool(g % 2 == 1 or left))])
    return rec


def h264_picture(seed, index, width_mbs, height_mbs, weighted=True, bi_fraction=0.4, deblock=True,
                 transform8_fraction=0.5, intra_fraction=0.0, mv_range=512, slices=1,
                 intra_types=(0.4, 0.25, 0.25, 0.1), transform_bypass=False, qp0_fraction=0.0,
                 scaling_lists=False):
    """One picture. References are supplied at fill time through extra pointers
    'ref{list}_{r}_{y|cb|cr}' (4 per list); POCs: list0 = [4, 2, 0, 6], list1 = [6, 8, 4, 2].
    intra_fraction of the macroblocks are intra, of the types Intra4x4 / Intra8x8 / Intra16x16 /
    I_PCM in the proportions intra_types; slices > 1 cuts the picture into that many raster
    slices (intra neighbours end at slice boundaries). transform_bypass sets
    qpprime_y_zero_transform_bypass_flag and qp0_fraction of the macroblocks get QP 0;
    scaling_lists gives the picture random scaling lists ([6][16] then [2][64], values 4..63)."""
    rng = np.random.Generator(np.random.PCG64([seed, 3, width_mbs, height_mbs, index]))
    n_mb = width_mbs * height_mbs
    mbs = np.zeros(n_mb, MB_DTYPE)
    blocks = np.zeros((n_mb, 16), BLOCK_DTYPE)
    mbs["qp"] = rng.integers(20, 41, n_mb)
    if qp0_fraction:
        mbs["qp"][rng.random(n_mb) < qp0_fraction] = 0
    t8 = rng.random(n_mb) < transform8_fraction
    intra = rng.random(n_mb) < intra_fraction
    mbs["flags"] = np.where(t8, abi.H264_MB_TRANSFORM_8X8, 0) | np.where(intra, abi.H264_MB_INTRA, 0)
    # partitions: 0 16x16, 1 16x8, 2 8x16, 3 8x8, 4 4x4 -> partition id of every 4x4 block
    ptype = rng.integers(0, 5, n_mb)
    bx, by = BLOCK_XY[:, 0], BLOCK_XY[:, 1]
    part_of = np.stack([np.zeros(16, int), by // 2, bx // 2, (by // 2) * 2 + bx // 2, np.arange(16)])
    pid = part_of[ptype]                                   # [n_mb, 16]
    nparts = 16
    bi = rng.random((n_mb, nparts)) < bi_fraction
    l1_only = rng.random((n_mb, nparts)) < 0.5
    ref0 = rng.integers(0, 4, (n_mb, nparts))
    ref1 = rng.integers(0, 4, (n_mb, nparts))
    r0 = np.where(bi | ~l1_only, ref0, -1)
    r1 = np.where(bi | l1_only, ref1, -1)
    mvs = rng.integers(-mv_range, mv_range + 1, (n_mb, nparts, 2, 2))
    take = lambda a: np.take_along_axis(a, pid.reshape(n_mb, 16, *([1] * (a.ndim - 2))), axis=1)
    blocks["ref_idx"][:, :, 0] = take(r0)
    blocks["ref_idx"][:, :, 1] = take(r1)
    blocks["mv"] = take(mvs)
    # residual
    cbp_luma = rng.integers(0, 16, n_mb)
    cbp_chroma = rng.integers(0, 3, n_mb)
    intra_rec = None
    record = np.zeros(n_mb, bool)   # level record present without a coded block pattern
    if intra.any() or slices > 1:
        if slices > 1:
            cuts = np.sort(rng.choice(np.arange(1, n_mb), slices - 1, replace=False))
            mbs["slice_id"] = np.searchsorted(cuts, np.arange(n_mb), side="right")
        intra_rec = _intra_records(rng, intra, mbs["slice_id"], width_mbs, np.asarray(intra_types) / sum(intra_types))
        itype = np.where(intra, intra_rec["type"], 255)
        t8 = np.where(intra, itype == abi.H264_INTRA_8X8, t8)
        mbs["flags"] = np.where(t8, abi.H264_MB_TRANSFORM_8X8, 0) | np.where(intra, abi.H264_MB_INTRA, 0)
        i16 = itype == abi.H264_INTRA_16X16
        pcm = itype == abi.H264_INTRA_PCM
        cbp_luma = np.where(i16, np.where(cbp_luma >= 8, 15, 0), np.where(pcm, 0, cbp_luma))
        cbp_chroma = np.where(pcm, 0, cbp_chroma)
        record = i16 | pcm
    cbp = (cbp_luma | (cbp_chroma << 4)).astype(np.uint8)
    mbs["cbp"] = cbp
    coded = np.nonzero((cbp != 0) | record)[0]
    mbs["coef_offset"][coded] = np.arange(coded.size)
    levels = np.zeros((max(coded.size, 1), abi.H264_LEVELS_PER_MB), np.int16)
    if coded.size:
        nc = coded.size
        luma = laplace_levels(rng, nc * 16, scale0=4.0, decay=3.0, zero_fraction=0.8, first=0, length=16).reshape(nc, 16, 16)
        lmask = ((cbp_luma[coded][:, None] >> (np.arange(16) // 4)) & 1).astype(bool)   # 8x8 group of each 4x4
        luma *= lmask[:, :, None]
        levels[:nc, :256] = luma.reshape(nc, 256)
        cdc = laplace_levels(rng, nc * 2, scale0=6.0, decay=100.0, zero_fraction=0.5, first=0, length=4).reshape(nc, 8)
        cac = laplace_levels(rng, nc * 8, scale0=3.0, decay=4.0, zero_fraction=0.8, first=1, length=15).reshape(nc, 120)
        cdc *= (cbp_chroma[coded] >= 1)[:, None]
        cac *= (cbp_chroma[coded] == 2)[:, None]
        levels[:nc, 256:264] = cdc
        levels[:nc, 264:] = cac
        if intra_rec is not None:
            ci = intra_rec["type"][coded]
            ii = intra[coded]
            i16 = ii & (ci == abi.H264_INTRA_16X16)
            if i16.any():   # luma DC levels by scan index at [16 k]
                dcl = laplace_levels(rng, int(i16.sum()), scale0=12.0, decay=4.0, zero_fraction=0.3, first=0, length=16)
                levels[:nc][i16, 0:256:16] = dcl
            pcm = ii & (ci == abi.H264_INTRA_PCM)
            if pcm.any():
                levels[:nc][pcm] = rng.integers(0, 256, (int(pcm.sum()), 384))
                luma[pcm] = 16          # counted as 16 nonzero coefficients per 4x4 (mark_pcm)
        # nonzero counts for bS: CAVLC total_coeff per 4x4; 8x8 transform: count of the 8x8 in all four
        nz4 = (luma != 0).sum(axis=2)
        t8c = t8[coded]
        nz8 = nz4.reshape(nc, 4, 4).sum(axis=2)
        nz = np.where(t8c[:, None], np.repeat(nz8, 4, axis=1), nz4)
        blocks["total_coeff"][coded] = np.minimum(nz, 255)
    weights = np.zeros((2, abi.H264_MAX_REFS, 3), WEIGHT_DTYPE)
    bi_w = np.zeros((abi.H264_MAX_REFS, abi.H264_MAX_REFS, 3), BI_WEIGHT_DTYPE)
    mode = abi.H264_WEIGHTED_NONE
    if weighted:
        mode = abi.H264_WEIGHTED_TABLE
        weights["weight"] = rng.integers(-64, 128, weights.shape)
        weights["offset"] = rng.integers(-16, 17, weights.shape)
        weights["log2_denom"] = 5
        for a in range(4):
            for b in range(4):
                for c in range(3):
                    bi_w[a, b, c] = (weights["weight"][0, a, c], weights["weight"][1, b, c],
                                     weights["offset"][0, a, c], weights["offset"][1, b, c], 5, (0, 0, 0))
    W, H = width_mbs * 16, height_mbs * 16
    outputs = {"y": ((H, W), np.uint8, None), "cb": ((H // 2, W // 2), np.uint8, None),
               "cr": ((H // 2, W // 2), np.uint8, None)}
    pocs = ([4, 2, 0, 6], [6, 8, 4, 2])

    def fill(p):
        s = abi.new(abi.H264Picture, width_mbs=width_mbs, height_mbs=height_mbs, weighted_mode=mode,
                    deblock=int(deblock), macroblocks=p["mbs"], blocks=p["blocks"], levels=p["levels"],
                    weights=p["weights"], bi_weights=p["bi_weights"], scaling_lists=p.get("scaling"),
                    intra=p.get("intra"), transform_bypass=int(transform_bypass))
        s.planes[0], s.planes[1], s.planes[2] = p["y"], p["cb"], p["cr"]
        s.chroma_qp_offset[0] = s.chroma_qp_offset[1] = 0
        for lst in range(2):
            s.ref_count[lst] = 4
            for r in range(4):
                ref = s.refs[lst][r]
                ref.poc = pocs[lst][r]
                for k, n in enumerate(("y", "cb", "cr")):
                    ref.planes[k] = p[f"ref{lst}_{r}_{n}"]
        return s

    ref_ids = ([0, 1, 2, 3], [3, 0, 1, 2])     # list -> resident reference index
    inputs = {"mbs": mbs, "blocks": blocks, "levels": levels, "weights": weights, "bi_weights": bi_w}
    if intra_rec is not None and intra.any():
        inputs["intra"] = intra_rec
    if scaling_lists:   # drawn last, so the other draws do not depend on it
        inputs["scaling"] = rng.integers(4, 64, 6 * 16 + 2 * 64, dtype=np.uint8)
    return Unit("h264", inputs,
                outputs, fill, {"width_mbs": width_mbs, "height_mbs": height_mbs, "ref_ids": ref_ids,
                                "coded_mbs": int(coded.size), "intra_mbs": int(intra.sum())})


def h264_ref_ptrs(unit, refs, kind):
    """Reference pointer entries from resident references: refs[i] = dict y/cb/cr of arrays (host)
    or device pointers."""
    extra = {}
    for lst in range(2):
        for r, rid in enumerate(unit.meta["ref_ids"][lst]):
            for n in ("y", "cb", "cr"):
                v = refs[rid][n]
                extra[f"ref{lst}_{r}_{n}"] = v.ctypes.data if kind == "host" else v
    return extra
