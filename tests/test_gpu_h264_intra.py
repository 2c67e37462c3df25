"""H.264 intra macroblocks on the device, bit-exact against the oracle: all-intra pictures (one
and several slices), inter pictures with intra macroblocks of every type, several pictures per
submission and one launch per picture, deblocking on and off, graph replays; the reference's
Intra16x16 known answers; intra data errors."""
import numpy as np
import pytest

import mmrecode_amd
from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

from test_gpu_parity import _assert_same, _h264_refs, _oracle_run
from test_h264_intra import KAT_EXPECT, intra16_kat_picture

pytestmark = pytest.mark.gpu


def _run(ctx, oracle, units, host_refs, dev_refs, flags=0):
    dptrs = [{n: b.ptr for n, b in r.items()} for r in dev_refs]
    dev = [wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, dptrs, "dev")) for u in units]
    try:
        ctx.wait(ctx.submit_h264([d.struct for d in dev], flags))
        for i, (u, d) in enumerate(zip(units, dev)):
            want = _oracle_run(oracle, "oracle_h264_reconstruct", u, wl.h264_ref_ptrs(u, host_refs, "host"))
            for n in ("y", "cb", "cr"):
                _assert_same(f"picture {i} {n}", d.download(n), want[n])
    finally:
        for d in dev:
            d.free()


@pytest.mark.parametrize("case", [
    dict(wmb=6, hmb=5, intra=1.0, slices=1, deblock=False),
    dict(wmb=9, hmb=7, intra=1.0, slices=3, deblock=True),
    dict(wmb=12, hmb=10, intra=0.3, slices=2, deblock=True),
    # many rows in flight: the row-to-row wavefront over 34 rows, rows ending at the right edge
    dict(wmb=23, hmb=34, intra=1.0, slices=1, deblock=True),
    dict(wmb=23, hmb=34, intra=0.5, slices=4, deblock=False, flags=abi.SUBMIT_PER_FRAME_LAUNCH),
    # only Intra4x4 / only Intra8x8 / only Intra16x16 + I_PCM
    dict(wmb=8, hmb=6, intra=1.0, slices=1, deblock=True, types=(1, 0, 0, 0)),
    dict(wmb=8, hmb=6, intra=1.0, slices=1, deblock=True, types=(0, 1, 0, 0)),
    dict(wmb=8, hmb=6, intra=1.0, slices=2, deblock=True, types=(0, 0, 3, 1)),
])
def test_h264_intra_pictures_bit_exact(ctx, oracle, case):
    wmb, hmb = case["wmb"], case["hmb"]
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    units = [wl.h264_picture(seed=51, index=i, width_mbs=wmb, height_mbs=hmb, intra_fraction=case["intra"],
                             slices=case["slices"], deblock=case["deblock"],
                             intra_types=case.get("types", (0.4, 0.25, 0.25, 0.1))) for i in range(3)]
    # a picture without intra macroblocks in the same submission
    units.append(wl.h264_picture(seed=52, index=0, width_mbs=wmb, height_mbs=hmb, deblock=case["deblock"]))
    try:
        _run(ctx, oracle, units, host_refs, dev_refs, case.get("flags", 0))
    finally:
        for r in dev_refs:
            for b in r.values():
                b.free()


@pytest.mark.parametrize("case", sorted(KAT_EXPECT))
def test_h264_intra16_reference_kats_on_device(ctx, case):
    mbs, blocks, levels, intra = intra16_kat_picture(case)
    bufs = {n: ctx.upload(a) for n, a in (("mbs", mbs), ("blocks", blocks), ("levels", levels), ("intra", intra))}
    planes = {"y": ctx.alloc(256), "cb": ctx.alloc(64), "cr": ctx.alloc(64)}
    try:
        pic = abi.new(abi.H264Picture, width_mbs=1, height_mbs=1, deblock=0, macroblocks=bufs["mbs"].ptr,
                      blocks=bufs["blocks"].ptr, levels=bufs["levels"].ptr, intra=bufs["intra"].ptr)
        pic.planes[0], pic.planes[1], pic.planes[2] = planes["y"].ptr, planes["cb"].ptr, planes["cr"].ptr
        pic.ref_count[0] = 1
        for k, n in enumerate(("y", "cb", "cr")):
            pic.refs[0][0].planes[k] = planes[n].ptr
        ctx.wait(ctx.submit_h264([pic]))
        out = {"y": ctx.download(planes["y"], np.uint8, (16, 16)), "cb": ctx.download(planes["cb"], np.uint8, (8, 8)),
               "cr": ctx.download(planes["cr"], np.uint8, (8, 8))}
        for plane, want in zip(("y", "cb", "cr"), KAT_EXPECT[case]):
            if want is None:
                continue
            if isinstance(want, tuple):
                assert tuple(out[plane].reshape(-1)[:4]) == want, (case, plane)
            else:
                assert (out[plane] == want).all(), (case, plane)
    finally:
        for b in list(bufs.values()) + list(planes.values()):
            b.free()


@pytest.mark.parametrize("what", ["mode_needs_missing_top", "bad_type", "8x8_without_transform_flag"])
def test_h264_intra_invalid_data(ctx, what):
    wmb, hmb = 4, 3
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    u = wl.h264_picture(seed=53, index=0, width_mbs=wmb, height_mbs=hmb, intra_fraction=1.0, deblock=False)
    rec = u.inputs["intra"]
    if what == "mode_needs_missing_top":
        rec["type"][0], rec["modes"][0, 0] = abi.H264_INTRA_4X4, 0      # vertical in the top-left macroblock
        u.inputs["mbs"]["flags"][0] &= ~np.uint8(abi.H264_MB_TRANSFORM_8X8)
    elif what == "bad_type":
        rec["type"][5] = 7
    else:
        rec["type"][5] = abi.H264_INTRA_8X8
        rec["modes"][5, :4] = 2
        u.inputs["mbs"]["flags"][5] &= ~np.uint8(abi.H264_MB_TRANSFORM_8X8)
    dptrs = [{n: b.ptr for n, b in r.items()} for r in dev_refs]
    d = wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, dptrs, "dev"))
    try:
        with pytest.raises(mmrecode_amd.ReconError) as err:
            ctx.wait(ctx.submit_h264([d.struct]))
        assert err.value.status == abi.STATUS_INVALID_DATA
        ctx.synchronize()
    finally:
        d.free()
        for r in dev_refs:
            for b in r.values():
                b.free()


def test_h264_intra_without_records_is_unsupported(ctx):
    wmb, hmb = 4, 3
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    u = wl.h264_picture(seed=54, index=0, width_mbs=wmb, height_mbs=hmb, intra_fraction=0.5, deblock=False)
    del u.inputs["intra"]
    dptrs = [{n: b.ptr for n, b in r.items()} for r in dev_refs]
    d = wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, dptrs, "dev"))
    try:
        with pytest.raises(mmrecode_amd.ReconError) as err:
            ctx.wait(ctx.submit_h264([d.struct]))
        assert err.value.status == abi.STATUS_UNSUPPORTED
        ctx.synchronize()
    finally:
        d.free()
        for r in dev_refs:
            for b in r.values():
                b.free()


def test_h264_intra_prepared_replays(ctx, oracle):
    """Row tickets and progress counters restart at zero on every replay of the captured graph."""
    wmb, hmb = 10, 12
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    units = [wl.h264_picture(seed=55, index=i, width_mbs=wmb, height_mbs=hmb, intra_fraction=0.6, slices=2, deblock=True)
             for i in range(2)]
    dptrs = [{n: b.ptr for n, b in r.items()} for r in dev_refs]
    dev = [wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, dptrs, "dev")) for u in units]
    want = [_oracle_run(oracle, "oracle_h264_reconstruct", u, wl.h264_ref_ptrs(u, host_refs, "host")) for u in units]
    prep = ctx.prepare(abi.CODEC_H264, [d.struct for d in dev])
    try:
        for r in range(3):
            for d in dev:
                for n in ("y", "cb", "cr"):
                    shape, dtype, _ = d.unit.outputs[n]
                    ctx.write(d.bufs[n], np.full(shape, 0x5A, dtype))
            ctx.wait(prep.launch())
            for i, d in enumerate(dev):
                for n in ("y", "cb", "cr"):
                    _assert_same(f"replay {r} picture {i} {n}", d.download(n), want[i][n])
            ctx.synchronize()
    finally:
        prep.close()
        for d in dev:
            d.free()
        for rr in dev_refs:
            for b in rr.values():
                b.free()
