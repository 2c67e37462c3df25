"""H.264 transform bypass on the device: the lossless round-trip pictures reconstruct to their
targets (inter 4x4 luma and chroma; Intra4x4 with the DPCM of vertical / horizontal modes and the
chroma continuation), and random pictures with QP 0 macroblocks of every type, with and without
the flag, are bit-exact against the oracle. Also pictures with their own scaling lists (every
other test uses Flat_16)."""
import ctypes as C

import numpy as np
import pytest

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

from test_gpu_parity import _assert_same, _h264_refs, _oracle_run
from test_h264_bypass import lossless_inter, lossless_intra

pytestmark = pytest.mark.gpu


def _device_picture(ctx, wmb, hmb, mbs, blocks, levels, intra=None, ref=None):
    bufs = {n: ctx.upload(a) for n, a in (("mbs", mbs), ("blocks", blocks), ("levels", levels))}
    if intra is not None:
        bufs["intra"] = ctx.upload(intra)
    planes = {"y": ctx.alloc(256 * wmb * hmb), "cb": ctx.alloc(64 * wmb * hmb), "cr": ctx.alloc(64 * wmb * hmb)}
    refs = {n: ctx.upload(a) for n, a in ref.items()} if ref is not None else planes
    pic = abi.new(abi.H264Picture, width_mbs=wmb, height_mbs=hmb, deblock=0, macroblocks=bufs["mbs"].ptr,
                  blocks=bufs["blocks"].ptr, levels=bufs["levels"].ptr,
                  intra=bufs["intra"].ptr if intra is not None else None, transform_bypass=1)
    for k, n in enumerate(("y", "cb", "cr")):
        pic.planes[k] = planes[n].ptr
        pic.refs[0][0].planes[k] = refs[n].ptr
    pic.ref_count[0] = 1
    allb = list(bufs.values()) + list(planes.values()) + (list(refs.values()) if ref is not None else [])
    return pic, planes, allb


@pytest.mark.parametrize("kind", ["inter", "intra"])
def test_lossless_round_trip_on_device(ctx, oracle, kind):
    rng = np.random.default_rng(17)
    wmb, hmb = 5, 4
    if kind == "inter":
        mbs, blocks, levels, ref, target = lossless_inter(rng, wmb, hmb)
        intra = None
    else:
        mbs, blocks, levels, intra, target = lossless_intra(oracle, rng, wmb, hmb)
        ref = None
    pic, planes, bufs = _device_picture(ctx, wmb, hmb, mbs, blocks, levels, intra, ref)
    try:
        ctx.wait(ctx.submit_h264([pic]))
        for n, shape in (("y", (16 * hmb, 16 * wmb)), ("cb", (8 * hmb, 8 * wmb)), ("cr", (8 * hmb, 8 * wmb))):
            _assert_same(f"{kind} {n}", ctx.download(planes[n], np.uint8, shape), target[n])
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("case", [
    dict(wmb=9, hmb=7, intra=0.5, flag=True, deblock=True),
    dict(wmb=9, hmb=7, intra=1.0, flag=True, deblock=False, types=(1, 0, 0, 0)),
    dict(wmb=12, hmb=5, intra=0.0, flag=True, deblock=True),
    dict(wmb=9, hmb=7, intra=0.5, flag=False, deblock=True),
    dict(wmb=9, hmb=7, intra=0.5, flag=False, deblock=True, scaling=True, qp0=0.0),
    dict(wmb=9, hmb=7, intra=1.0, flag=True, deblock=True, scaling=True),
])
def test_bypass_pictures_bit_exact(ctx, oracle, case):
    wmb, hmb = case["wmb"], case["hmb"]
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    units = [wl.h264_picture(seed=61, index=i, width_mbs=wmb, height_mbs=hmb, intra_fraction=case["intra"],
                             deblock=case["deblock"], intra_types=case.get("types", (0.4, 0.25, 0.25, 0.1)),
                             transform_bypass=case["flag"], qp0_fraction=case.get("qp0", 0.5),
                             scaling_lists=case.get("scaling", False)) for i in range(3)]
    dptrs = [{n: b.ptr for n, b in r.items()} for r in dev_refs]
    dev = [wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, dptrs, "dev")) for u in units]
    try:
        ctx.wait(ctx.submit_h264([d.struct for d in dev]))
        for i, (u, d) in enumerate(zip(units, dev)):
            want = _oracle_run(oracle, "oracle_h264_reconstruct", u, wl.h264_ref_ptrs(u, host_refs, "host"))
            for n in ("y", "cb", "cr"):
                _assert_same(f"picture {i} {n}", d.download(n), want[n])
    finally:
        for d in dev:
            d.free()
        for r in dev_refs:
            for b in r.values():
                b.free()
