"""The YUV CompositionGraph (crates/render/src/frame_graph.rs:234-283, executed by the CPU backend at
:370-430) end to end on the device: an MPEG-2 picture of the reference's progressive stream is
reconstructed, the Scale pass maps it onto the canvas (scale_yuv420_to_canvas_cpu), the
ColorConvert pass's prepared overlay planes come from overlay preparation (prepare_yuv) and the
Composite pass blends them (blend_yuv420) -- every intermediate stays in device memory, and the
canvas equals the oracle's chain bit for bit."""
import ctypes as C

import numpy as np
import pytest

import mpeg2_streams as ms
from mmrecode_amd import abi
from test_compose_scale import oracle_prepare, oracle_scale, overlay_shapes, random_overlay, scale_job
from test_gpu_parity import _assert_same, _mpeg2_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("canvas,mode,filt", [((160, 90), abi.SCALE_FIT, abi.SCALE_FILTER_LANCZOS3),
                                              ((64, 48), abi.SCALE_FILL, abi.SCALE_FILTER_TRIANGLE)])
def test_yuv_composition_graph_on_device(ctx, oracle, canvas, mode, filt):
    from oracle import binding

    cw, ch = canvas
    pics = binding.mpeg2_pictures(ms.read("main-ml-progressive-ibp.m2v"))
    units = ms.units(pics)
    want_frames = ms.oracle_decode(oracle, units)
    frame = max(range(len(pics)), key=lambda i: pics[i]["display_index"] == 5)   # a B picture
    dev = _mpeg2_gpu(ctx, units)
    bufs = []
    try:
        # Scale: the decoded picture's device planes straight into the canvas
        canvas_planes = [ctx.alloc(cw * ch), ctx.alloc(cw * ch // 4), ctx.alloc(cw * ch // 4)]
        bufs += canvas_planes
        d = dev[frame]
        src = [d.ptr("y"), d.ptr("cb"), d.ptr("cr")]
        job = scale_job(src, [96, 48, 48], 96, 64, [b.ptr for b in canvas_planes], cw, ch, mode, abi.RANGE_LIMITED, filt)
        ctx.wait(ctx.submit_scale([job]))

        # ColorConvert: a straight-alpha overlay prepared into its YUV planes
        rng = np.random.default_rng(cw + ch)
        ow, oh, ox, oy = 40, 30, 12, 8
        rgba = random_overlay(rng, ow, oh)
        d_rgba = ctx.upload(rgba)
        shapes = overlay_shapes(ow, oh)
        prep = {n: ctx.alloc(int(np.prod(s))) for n, s in shapes.items()}
        bufs += [d_rgba] + list(prep.values())
        ctx.wait(ctx.submit_overlay_prepare([abi.new(abi.OverlayPrepareJob, rgba=d_rgba.ptr, width=ow, height=oh,
                                                     **{n: b.ptr for n, b in prep.items()})]))

        # Composite: the limited-range planes blended over the canvas at (ox, oy)
        blend = abi.new(abi.YuvBlendJob, target_width=cw, target_height=ch, rgba=d_rgba.ptr,
                        chroma_alpha=prep["chroma_alpha"].ptr, width=ow, height=oh, x=ox, y=oy)
        for k, n in enumerate(("y_limited", "u_limited", "v_limited")):
            blend.target[k] = canvas_planes[k].ptr
            blend.overlay[k] = prep[n].ptr
            blend.target_stride[k] = cw if k == 0 else cw // 2
        ctx.wait(ctx.submit_yuv_blend([blend]))
        got = [ctx.download(canvas_planes[0], np.uint8, (ch, cw)), ctx.download(canvas_planes[1], np.uint8, (ch // 2, cw // 2)),
               ctx.download(canvas_planes[2], np.uint8, (ch // 2, cw // 2))]

        # The oracle's chain on the host
        f = want_frames[frame]
        want = oracle_scale(oracle, [f["y"], f["cb"], f["cr"]], cw, ch, mode, abi.RANGE_LIMITED, filt)
        pw = oracle_prepare(oracle, rgba)
        hj = abi.new(abi.YuvBlendJob, target_width=cw, target_height=ch, rgba=rgba.ctypes.data,
                     chroma_alpha=pw["chroma_alpha"].ctypes.data, width=ow, height=oh, x=ox, y=oy)
        for k, n in enumerate(("y_limited", "u_limited", "v_limited")):
            hj.target[k] = want[k].ctypes.data
            hj.overlay[k] = pw[n].ctypes.data
            hj.target_stride[k] = cw if k == 0 else cw // 2
        assert oracle.oracle_yuv_blend(C.byref(hj)) == 0
        for k in range(3):
            _assert_same(f"canvas plane {k}", got[k], want[k])
    finally:
        for b in bufs:
            b.free()
        for x in dev:
            x.free()
