"""Overlay preparation and the 4:2:0 canvas scaler on the device, bit-exact against the oracle:
odd overlay sizes and absent outputs, every scale mode with both filters, up- and down-scaling,
strided sources, several jobs per submission and one launch per job; the reference's known
answers; argument errors."""
import numpy as np
import pytest

import mmrecode_amd
from mmrecode_amd import abi

from test_compose_scale import (OVERLAY_OUTPUTS, SCALE_CASES, oracle_prepare, oracle_scale, overlay_shapes,
                                random_overlay, scale_job, scale_planes)
from test_gpu_parity import _assert_same

pytestmark = pytest.mark.gpu


def _free(bufs):
    for b in bufs:
        b.free()


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_overlay_prepare_bit_exact(ctx, oracle, flags):
    rng = np.random.default_rng(31)
    sizes = [(8, 6), (7, 5), (1, 1), (33, 17), (2, 3), (1280, 720), (641, 359)]
    overlays = [random_overlay(rng, w, h) for w, h in sizes]
    bufs, jobs, outs = [], [], []
    try:
        for i, ((w, h), rgba) in enumerate(zip(sizes, overlays)):
            d_rgba = ctx.upload(rgba)
            bufs.append(d_rgba)
            shapes = overlay_shapes(w, h)
            skip = OVERLAY_OUTPUTS[i % len(OVERLAY_OUTPUTS)] if i % 3 == 2 else None   # one absent output
            o = {n: ctx.alloc(int(np.prod(s))) for n, s in shapes.items() if n != skip}
            bufs.extend(o.values())
            outs.append((o, shapes, skip))
            jobs.append(abi.new(abi.OverlayPrepareJob, rgba=d_rgba.ptr, width=w, height=h, **{n: b.ptr for n, b in o.items()}))
        ctx.wait(ctx.submit_overlay_prepare(jobs, flags))
        for i, rgba in enumerate(overlays):
            want = oracle_prepare(oracle, rgba)
            o, shapes, skip = outs[i]
            for n, b in o.items():
                _assert_same(f"overlay {i} {n}", ctx.download(b, np.uint8, shapes[n]), want[n])
    finally:
        _free(bufs)


def _device_scale(ctx, planes, cw, ch, mode, rng_range, filt, pad=0):
    """Uploads the planes with `pad` extra bytes per row (a stride above the width)."""
    src, strides = [], []
    for p in planes:
        wide = np.zeros((p.shape[0], p.shape[1] + pad), np.uint8)
        wide[:, :p.shape[1]] = p
        src.append(ctx.upload(wide))
        strides.append(wide.shape[1])
    dst = [ctx.alloc(ch * cw), ctx.alloc(ch * cw // 4), ctx.alloc(ch * cw // 4)]
    h, w = planes[0].shape
    job = scale_job([b.ptr for b in src], strides, w, h, [b.ptr for b in dst], cw, ch, mode, rng_range, filt)
    return job, src, dst


def _download_canvas(ctx, dst, cw, ch):
    return [ctx.download(dst[0], np.uint8, (ch, cw)), ctx.download(dst[1], np.uint8, (ch // 2, cw // 2)),
            ctx.download(dst[2], np.uint8, (ch // 2, cw // 2))]


BIG_CASES = [
    (1920, 1080, 1280, 720, abi.SCALE_FIT),      # 1080p preview downscale
    (720, 576, 1920, 1080, abi.SCALE_FIT),       # SD onto an HD canvas: upscale with pillar bars
    (1280, 720, 1080, 1920, abi.SCALE_FILL),     # landscape cropped into portrait
    (640, 480, 1920, 1080, abi.SCALE_STRETCH),
]


@pytest.mark.parametrize("filt", [abi.SCALE_FILTER_LANCZOS3, abi.SCALE_FILTER_TRIANGLE])
@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_scale_bit_exact(ctx, oracle, filt, flags):
    rng = np.random.default_rng(41 + filt)
    cases = SCALE_CASES + BIG_CASES
    bufs, jobs, wants = [], [], []
    try:
        for i, (w, h, cw, ch, mode) in enumerate(cases):
            planes = scale_planes(rng, w, h, "random" if i % 2 else "smooth")
            rng_range = abi.RANGE_LIMITED if i % 2 else abi.RANGE_FULL
            job, src, dst = _device_scale(ctx, planes, cw, ch, mode, rng_range, filt, pad=(i % 3) * 24)
            bufs.extend(src + dst)
            jobs.append((job, dst, cw, ch))
            wants.append(oracle_scale(oracle, planes, cw, ch, mode, rng_range, filt))
        ctx.wait(ctx.submit_scale([j for j, _, _, _ in jobs], flags))
        for i, ((_, dst, cw, ch), want) in enumerate(zip(jobs, wants)):
            got = _download_canvas(ctx, dst, cw, ch)
            for k in range(3):
                _assert_same(f"case {cases[i]} plane {k}", got[k], want[k])
    finally:
        _free(bufs)


def test_scale_reference_kats_on_device(ctx):
    """compositor.rs:351-369 on the device: fit, stretch and native of the constant 4x4 frame."""
    frame = scale_planes(None, 4, 4, "constant")
    for mode in (abi.SCALE_FIT, abi.SCALE_STRETCH, abi.SCALE_NATIVE):
        job, src, dst = _device_scale(ctx, frame, 8, 4, mode, abi.RANGE_LIMITED, abi.SCALE_FILTER_LANCZOS3)
        try:
            ctx.wait(ctx.submit_scale([job]))
            y, u, _ = _download_canvas(ctx, dst, 8, 4)
            if mode == abi.SCALE_STRETCH:
                assert (y == 200).all()
            else:
                assert list(y[0]) == [16, 16, 200, 200, 200, 200, 16, 16]
                assert u[0, 0] == 128
        finally:
            _free(src + dst)


def test_scale_and_prepare_reject_bad_arguments(ctx):
    frame = scale_planes(None, 4, 4, "constant")
    for field, value, status in (("width", 5, abi.STATUS_INVALID_DATA), ("canvas_height", 0, abi.STATUS_INVALID_DATA),
                                 ("mode", 9, abi.STATUS_INVALID_ARGUMENT), ("filter", 4, abi.STATUS_INVALID_ARGUMENT)):
        job, src, dst = _device_scale(ctx, frame, 8, 4, abi.SCALE_FIT, abi.RANGE_LIMITED, abi.SCALE_FILTER_LANCZOS3)
        try:
            setattr(job, field, value)
            with pytest.raises(mmrecode_amd.ReconError) as err:
                ctx.submit_scale([job])
            assert err.value.status == status, field
        finally:
            _free(src + dst)
    rgba = ctx.alloc(64)
    try:
        job = abi.new(abi.OverlayPrepareJob, rgba=rgba.ptr + 1, width=2, height=2)
        with pytest.raises(mmrecode_amd.ReconError) as err:
            ctx.submit_overlay_prepare([job])
        assert err.value.status == abi.STATUS_INVALID_ARGUMENT
    finally:
        rgba.free()
    ctx.synchronize()
