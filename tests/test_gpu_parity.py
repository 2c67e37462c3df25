"""Bit-exact parity of the HIP path (through the C ABI) against the oracle on seeded inputs."""
import ctypes as C

import numpy as np
import pytest

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, fn_name, unit, extra=None):
    out = unit.host_outputs()
    s = unit.host_struct(out, extra)
    rc = getattr(oracle, fn_name)(C.byref(s))
    assert rc == 0, f"{fn_name} returned {rc}"
    return out


def _assert_same(name, got, want):
    if not np.array_equal(got, want):
        diff = np.argwhere(got != want)
        first = tuple(diff[0])
        raise AssertionError(f"{name}: {len(diff)} samples differ; first at {first}: "
                             f"gpu {got[first]} oracle {want[first]}")


# ------------------------------------------------------------------ MJPEG
@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_mjpeg_config0_frames_bit_exact(ctx, oracle, flags):
    units = [wl.mjpeg_frame(seed=11, index=i) for i in range(3)]
    dev = [wl.DeviceUnit(ctx, u) for u in units]
    ticket = ctx.submit_mjpeg([d.struct for d in dev], flags)
    ctx.wait(ticket)
    for u, d in zip(units, dev):
        want = _oracle_run(oracle, "oracle_mjpeg_reconstruct", u)
        for name in want:
            _assert_same(name, d.download(name), want[name])
        d.free()


@pytest.mark.parametrize("geometry", [
    (17, 13, ((2, 2), (1, 1), (1, 1))),     # ragged 4:2:0, partial MCUs
    (33, 9, ((2, 1), (1, 1), (1, 1))),      # 4:2:2
    (24, 24, ((1, 1), (1, 1), (1, 1))),     # 4:4:4
    (21, 11, ((1, 1),)),                     # grayscale, non-interleaved
])
def test_mjpeg_sampling_and_edges(ctx, oracle, geometry):
    w, h, sampling = geometry
    u = wl.mjpeg_frame(seed=5, index=w * 100 + h, width=w, height=h, sampling=sampling)
    d = wl.DeviceUnit(ctx, u)
    ctx.wait(ctx.submit_mjpeg([d.struct]))
    want = _oracle_run(oracle, "oracle_mjpeg_reconstruct", u)
    for name in want:
        _assert_same(name, d.download(name), want[name])
    d.free()


# ------------------------------------------------------------------ MPEG-2
def _mpeg2_gpu(ctx, units, flags=0):
    dev = []
    for u in units:
        extra = {"forward_source": -1 if u.meta["fwd"] is None else u.meta["fwd"],
                 "backward_source": -1 if u.meta["bwd"] is None else u.meta["bwd"]}
        for tag in ("fwd", "bwd"):
            ref = u.meta[tag]
            if ref is not None:
                for n in ("y", "cb", "cr"):
                    extra[f"{tag}_{n}"] = dev[ref].ptr(n)
        dev.append(wl.DeviceUnit(ctx, u, extra))
    ctx.wait(ctx.submit_mpeg2([d.struct for d in dev], flags))
    return dev


def _mpeg2_oracle(oracle, units):
    outs = {}
    for i, u in enumerate(units):
        outs[i] = _oracle_run(oracle, "oracle_mpeg2_reconstruct", u, wl.mpeg2_ref_ptrs(outs, u, "host"))
    return outs


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
@pytest.mark.parametrize("size", [(352, 288), (720, 480), (100, 50)])
def test_mpeg2_gop_bit_exact(ctx, oracle, flags, size):
    units = wl.mpeg2_gop(seed=21, gop_index=size[0], width=size[0], height=size[1])
    dev = _mpeg2_gpu(ctx, units, flags)
    want = _mpeg2_oracle(oracle, units)
    for i, d in enumerate(dev):
        for n in ("y", "cb", "cr"):
            _assert_same(f"picture {i} ({units[i].meta['type']}) {n}", d.download(n), want[i][n])
    for d in dev:
        d.free()


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
@pytest.mark.parametrize("size", [(176, 144), (100, 50)])
def test_mpeg2_field_prediction_bit_exact(ctx, oracle, flags, size):
    """Field prediction (random field vectors reaching past the edges, random selects) and field
    DCT mixed with frame prediction."""
    units = wl.mpeg2_gop(seed=27, gop_index=size[1], width=size[0], height=size[1], field_fraction=0.5)
    assert any(((u.inputs["mbs"]["flags"] & abi.MPEG2_MB_FIELD_MOTION) != 0).any() for u in units)
    dev = _mpeg2_gpu(ctx, units, flags)
    want = _mpeg2_oracle(oracle, units)
    for i, d in enumerate(dev):
        for n in ("y", "cb", "cr"):
            _assert_same(f"picture {i} ({units[i].meta['type']}) {n}", d.download(n), want[i][n])
    for d in dev:
        d.free()


@pytest.mark.parametrize("case", ["no_field_array", "bad_select"])
def test_mpeg2_field_motion_invalid_data(ctx, case):
    import mmrecode_amd

    units = wl.mpeg2_gop(seed=3, width=64, height=32, gop="IP", field_fraction=0.5)
    field = (units[1].inputs["mbs"]["flags"] & abi.MPEG2_MB_FIELD_MOTION) != 0
    assert field.any()
    if case == "no_field_array":
        del units[1].inputs["field_motion"]
    else:
        units[1].inputs["field_motion"]["field_select"][np.argmax(field), 1, 0] = 2
    with pytest.raises(mmrecode_amd.ReconError) as err:
        _mpeg2_gpu(ctx, units)
    assert err.value.status == abi.STATUS_INVALID_DATA
    ctx.synchronize()   # the error was returned by wait and is not repeated


# ------------------------------------------------------------------ DV
@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_dv_mixed_systems_bit_exact(ctx, oracle, flags):
    systems = [abi.DV_SYSTEM_625_50, abi.DV_SYSTEM_525_60, abi.DV_SYSTEM_525_60, abi.DV_SYSTEM_625_50]
    units = [wl.dv_frame(seed=31, index=i, system=s) for i, s in enumerate(systems)]
    dev = [wl.DeviceUnit(ctx, u) for u in units]
    ctx.wait(ctx.submit_dv([d.struct for d in dev], flags))
    for i, (u, d) in enumerate(zip(units, dev)):
        want = _oracle_run(oracle, "oracle_dv_reconstruct", u)
        for n in ("y", "cb", "cr"):
            _assert_same(f"frame {i} {n}", d.download(n), want[n])
        d.free()


# ------------------------------------------------------------------ H.264
def _h264_refs(ctx, wmb, hmb, seed=7):
    host = [dict(zip(("y", "cb", "cr"), wl.h264_reference_planes(seed, wmb, hmb, i))) for i in range(4)]
    dev = [{n: ctx.upload(a) for n, a in r.items()} for r in host]
    return host, dev


@pytest.mark.parametrize("case", [
    dict(wmb=6, hmb=5, weighted=True, deblock=False),
    dict(wmb=6, hmb=5, weighted=False, deblock=False),
    dict(wmb=7, hmb=6, weighted=True, deblock=True),
    dict(wmb=9, hmb=4, weighted=False, deblock=True),
    # deblocking bands of 16 macroblock rows: 3 bands (16, 16, 8), hand-over between workgroups
    dict(wmb=12, hmb=40, weighted=True, deblock=True),
    # wide picture (ring slots reused), last band of a single row
    dict(wmb=44, hmb=33, weighted=False, deblock=True),
    # one launch per picture
    dict(wmb=10, hmb=34, weighted=True, deblock=True, flags=abi.SUBMIT_PER_FRAME_LAUNCH),
])
def test_h264_pictures_bit_exact(ctx, oracle, case):
    wmb, hmb = case["wmb"], case["hmb"]
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    units = [wl.h264_picture(seed=41, index=i, width_mbs=wmb, height_mbs=hmb, weighted=case["weighted"],
                             deblock=case["deblock"]) for i in range(3)]
    dev = [wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, [{n: b.ptr for n, b in r.items()} for r in dev_refs], "dev"))
           for u in units]
    ctx.wait(ctx.submit_h264([d.struct for d in dev], case.get("flags", 0)))
    for i, (u, d) in enumerate(zip(units, dev)):
        want = _oracle_run(oracle, "oracle_h264_reconstruct", u, wl.h264_ref_ptrs(u, host_refs, "host"))
        for n in ("y", "cb", "cr"):
            _assert_same(f"picture {i} {n}", d.download(n), want[n])
        d.free()
    for r in dev_refs:
        for b in r.values():
            b.free()


# ------------------------------------------------------------------ colour / composite
@pytest.mark.parametrize("fmt", [(64, 48, 32, 24), (37, 21, 19, 11), (720, 480, 180, 480)])
@pytest.mark.parametrize("rng_matrix", [(abi.RANGE_FULL, abi.MATRIX_BT601), (abi.RANGE_LIMITED, abi.MATRIX_BT601),
                                        (abi.RANGE_LIMITED, abi.MATRIX_BT709), (abi.RANGE_FULL, abi.MATRIX_BT709)])
def test_yuv_to_rgb_bit_exact(ctx, oracle, fmt, rng_matrix):
    w, h, cw, ch = fmt
    rng = np.random.default_rng(w * 7 + h)
    planes = [rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (ch, cw), dtype=np.uint8),
              rng.integers(0, 256, (ch, cw), dtype=np.uint8)]
    rgb = np.zeros((h, w, 3), np.uint8)
    rgba = np.zeros((h, w, 4), np.uint8)
    job = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=w,
                  chroma_stride=cw, range=rng_matrix[0], matrix=rng_matrix[1], rgb=rgb.ctypes.data, rgba=rgba.ctypes.data)
    for k in range(3):
        job.planes[k] = planes[k].ctypes.data
    assert oracle.oracle_yuv_to_rgb(C.byref(job)) == 0
    bufs = [ctx.upload(p) for p in planes]
    drgb, drgba = ctx.alloc(rgb.nbytes), ctx.alloc(rgba.nbytes)
    djob = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=w,
                   chroma_stride=cw, range=rng_matrix[0], matrix=rng_matrix[1], rgb=drgb.ptr, rgba=drgba.ptr)
    for k in range(3):
        djob.planes[k] = bufs[k].ptr
    ctx.wait(ctx.submit_yuv_to_rgb([djob]))
    _assert_same("rgb", ctx.download(drgb, np.uint8, rgb.shape), rgb)
    _assert_same("rgba", ctx.download(drgba, np.uint8, rgba.shape), rgba)
    for b in bufs + [drgb, drgba]:
        b.free()


def _random_rgba(rng, h, w, premultiplied=False):
    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    alpha = rng.choice([0, 255, 1, 128, 254], size=(h, w), p=[0.15, 0.15, 0.05, 0.05, 0.6])
    px[..., 3] = np.where(alpha == 254, rng.integers(0, 256, (h, w)), alpha)
    if premultiplied:
        px[..., :3] = (px[..., :3].astype(np.uint32) * px[..., 3:4] // 255).astype(np.uint8)
    return px


@pytest.mark.parametrize("op", [abi.BLEND_SOURCE_OVER, abi.BLEND_PREMULTIPLIED_OVER])
def test_composite_layers_bit_exact(ctx, oracle, op):
    rng = np.random.default_rng(op)
    H, W = 45, 67
    target = _random_rgba(rng, H, W)
    layers = [(_random_rgba(rng, 30, 40, op == 2), 3, 5), (_random_rgba(rng, 45, 67, op == 2), 0, 0),
              (_random_rgba(rng, 10, 9, op == 2), 50, 30)]
    host_target = target.copy()
    job = abi.new(abi.CompositeJob, target=host_target.ctypes.data, width=W, height=H, stride=W * 4, layer_count=3)
    for i, (px, x, y) in enumerate(layers):
        job.layers[i] = abi.Layer(rgba=px.ctypes.data, width=px.shape[1], height=px.shape[0], stride=px.shape[1] * 4,
                                  x=x, y=y, op=op, opacity=200 if op == 2 else 255)
    assert oracle.oracle_composite(C.byref(job)) == 0
    dt = ctx.upload(target)
    dl = [ctx.upload(px) for px, _, _ in layers]
    djob = abi.new(abi.CompositeJob, target=dt.ptr, width=W, height=H, stride=W * 4, layer_count=3)
    for i, ((px, x, y), b) in enumerate(zip(layers, dl)):
        djob.layers[i] = abi.Layer(rgba=b.ptr, width=px.shape[1], height=px.shape[0], stride=px.shape[1] * 4,
                                   x=x, y=y, op=op, opacity=200 if op == 2 else 255)
    ctx.wait(ctx.submit_composite([djob]))
    _assert_same("target", ctx.download(dt, np.uint8, target.shape), host_target)
    for b in [dt] + dl:
        b.free()


def test_yuv420_overlay_blend_bit_exact(ctx, oracle):
    rng = np.random.default_rng(9)
    TW, TH, w, h, x, y = 64, 48, 30, 20, 8, 6
    target = [rng.integers(0, 256, (TH, TW), dtype=np.uint8), rng.integers(0, 256, (TH // 2, TW // 2), dtype=np.uint8),
              rng.integers(0, 256, (TH // 2, TW // 2), dtype=np.uint8)]
    overlay = [rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
               rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)]
    rgba = _random_rgba(rng, h, w)
    calpha = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
    host = [t.copy() for t in target]

    def job_of(tp, op, ap, cp):
        j = abi.new(abi.YuvBlendJob, target_width=TW, target_height=TH, rgba=ap, chroma_alpha=cp, width=w, height=h, x=x, y=y)
        for k in range(3):
            j.target[k] = tp[k]
            j.overlay[k] = op[k]
            j.target_stride[k] = TW if k == 0 else TW // 2
        return j

    hj = job_of([t.ctypes.data for t in host], [o.ctypes.data for o in overlay], rgba.ctypes.data, calpha.ctypes.data)
    assert oracle.oracle_yuv_blend(C.byref(hj)) == 0
    dt = [ctx.upload(t) for t in target]
    do = [ctx.upload(o) for o in overlay]
    da, dc = ctx.upload(rgba), ctx.upload(calpha)
    ctx.wait(ctx.submit_yuv_blend([job_of([b.ptr for b in dt], [b.ptr for b in do], da.ptr, dc.ptr)]))
    for k in range(3):
        _assert_same(f"plane {k}", ctx.download(dt[k], np.uint8, target[k].shape), host[k])
    for b in dt + do + [da, dc]:
        b.free()
