"""Algorithmic byte counts and throughput units of the reconstruction workloads.

Algorithmic bytes = unique inputs that must be read (coefficient levels as dense int16, 128 B per
coded 8x8 block; per-block / per-macroblock records; reference samples at 1 B per predicted
sample per list; tables once) plus final outputs written once. Intermediate re-reads are not
counted. Mpix = visible luma pixels.
"""
import numpy as np

MPEG2_MB_RECORD = 16


def mpeg2_picture_bytes(unit):
    """Algorithmic bytes of one MPEG-2 picture Unit (mmrecode_amd.workloads.mpeg2_gop)."""
    W = -(-unit.meta["width"] // 16) * 16
    H = -(-unit.meta["height"] // 16) * 16
    samples = W * H * 3 // 2
    kinds = unit.meta["kinds"]                      # intra, fwd, bwd, bi macroblock counts
    ref_samples_per_mb = 384
    ref_bytes = (kinds[1] + kinds[2] + 2 * kinds[3]) * ref_samples_per_mb
    n_mb = int(kinds.sum())
    return (unit.meta["coded_blocks"] * 128 + n_mb * MPEG2_MB_RECORD + 4 * 64 + ref_bytes + samples)


def mjpeg_frame_bytes(unit):
    out = sum(int(np.prod(s)) for s, _, _ in unit.outputs.values())
    return unit.meta["blocks"] * 128 + 2 * 64 * 2 + out


def mjpeg_plane_bytes(unit):
    return sum(int(np.prod(s)) for n, (s, _, _) in unit.outputs.items() if n != "rgb")


def mjpeg_block_kernel_bytes(unit):
    """The block kernel's share: levels, quantisation tables, component planes written once."""
    return unit.meta["blocks"] * 128 + 2 * 64 * 2 + mjpeg_plane_bytes(unit)


def mjpeg_rgb_kernel_bytes(unit):
    """The YUV->RGB kernel's share: the three planes read once, packed RGB written once."""
    if "rgb" not in unit.outputs:
        return 0
    return mjpeg_plane_bytes(unit) + int(np.prod(unit.outputs["rgb"][0]))


def dv_frame_bytes(unit):
    out = sum(int(np.prod(s)) for s, _, _ in unit.outputs.values())
    return unit.meta["blocks"] * (128 + 1) + out


def h264_picture_bytes(unit):
    """Levels of coded macroblocks (384 x int16), 12 B per macroblock record, 12 B per 4x4 block
    record, reference reads of inter macroblocks (1 B per predicted sample per list used), 20 B
    per intra record, weights, output planes. Intra neighbours are re-reads of output samples."""
    mbs = unit.inputs["mbs"]
    blocks = unit.inputs["blocks"]
    n_mb = mbs.size
    inter = (mbs["flags"] & 1) == 0
    lists_used = (blocks["ref_idx"][:, :, 0] >= 0).astype(int) + (blocks["ref_idx"][:, :, 1] >= 0).astype(int)
    ref_bytes = int(lists_used[inter].sum()) * 24   # 16 luma + 2 x 4 chroma samples per 4x4 block per list
    out = sum(int(np.prod(s)) for s, _, _ in unit.outputs.values())
    intra_bytes = int((~inter).sum()) * 20
    return (unit.meta["coded_mbs"] * 768 + n_mb * 12 + n_mb * 16 * 12 + unit.inputs["weights"].nbytes
            + unit.inputs["bi_weights"].nbytes + ref_bytes + intra_bytes + out)


def mpix(width, height):
    return width * height / 1e6
