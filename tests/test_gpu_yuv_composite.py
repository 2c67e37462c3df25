"""Conversion and composite in one pass (mmr_recon_yuv_composite_submit) bit-exact against the
oracle's ycbcr_to_rgb then blend chain and against the two-pass path on the device: both ranges
and matrices, 4:2:0 / 4:2:2 / 4:1:1 chroma, widths that are not a multiple of 4 (unaligned rows,
per-pixel tails), layers placed partly inside the target, straight-alpha and premultiplied layers
at opacities below 255, several jobs per submission and one launch per job; a composite whose
target is not the conversion's packed output is refused."""
import ctypes as C

import numpy as np
import pytest

from mmrecode_amd import abi
from mmrecode_amd import ReconError

from test_gpu_parity import _assert_same

pytestmark = pytest.mark.gpu

# (width, height, chroma width, chroma height, range, matrix)
CASES = [(640, 480, 320, 240, abi.RANGE_LIMITED, abi.MATRIX_BT709),
         (720, 480, 180, 480, abi.RANGE_LIMITED, abi.MATRIX_BT601),     # 4:1:1
         (718, 37, 359, 19, abi.RANGE_FULL, abi.MATRIX_BT601),          # width % 4 == 2, odd chroma height
         (257, 9, 129, 5, abi.RANGE_FULL, abi.MATRIX_BT709),            # odd width: unaligned rows
         (1922, 33, 961, 33, abi.RANGE_LIMITED, abi.MATRIX_BT709),      # 4:2:2, 8 waves per row
         (3, 2, 2, 1, abi.RANGE_LIMITED, abi.MATRIX_BT601)]


def _layers(rng, w, h):
    """Three layers: a full-frame premultiplied one at opacity 255, a straight-alpha one placed
    inside, and a premultiplied one at the bottom-right corner at a random opacity."""
    out = []
    specs = [(w, h, 0, 0, abi.BLEND_PREMULTIPLIED_OVER, 255),
             (max(1, w // 2 + 1), max(1, h // 3 + 1), w // 5, h // 4, abi.BLEND_SOURCE_OVER, 255),
             (max(1, w // 3), max(1, h // 2), w - max(1, w // 3), h - max(1, h // 2), abi.BLEND_PREMULTIPLIED_OVER,
              int(rng.integers(1, 255)))]
    for lw, lh, x, y, op, opacity in specs:
        px = rng.integers(0, 256, (lh, lw, 4), dtype=np.uint8)
        if op == abi.BLEND_PREMULTIPLIED_OVER:
            px[..., :3] = (px[..., :3].astype(np.uint32) * px[..., 3:4] // 255).astype(np.uint8)
        out.append((px, x, y, op, opacity))
    return out


def _jobs(case, planes, rgba_ptr, layers, layer_ptrs, strides):
    w, h, cw, ch, rng_, mat = case
    conv = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=strides[0],
                   chroma_stride=strides[1], range=rng_, matrix=mat, rgba=rgba_ptr)
    for k in range(3):
        conv.planes[k] = planes[k]
    comp = abi.new(abi.CompositeJob, target=rgba_ptr, width=w, height=h, stride=w * 4, layer_count=len(layers))
    for i, ((px, x, y, op, opacity), p) in enumerate(zip(layers, layer_ptrs)):
        comp.layers[i] = abi.Layer(rgba=p, width=px.shape[1], height=px.shape[0], stride=px.shape[1] * 4, x=x, y=y, op=op,
                                   opacity=opacity)
    return conv, comp


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_yuv_composite_bit_exact(ctx, oracle, flags):
    rng = np.random.default_rng(404)
    keep, fused, two_pass, conv_d, comp_d, conv2_d, comp2_d, want = [], [], [], [], [], [], [], []
    try:
        for case in CASES:
            w, h, cw, ch = case[:4]
            strides = (w + 3, cw + 1)   # strided planes
            planes = [rng.integers(0, 256, (h, strides[0]), dtype=np.uint8), rng.integers(0, 256, (ch, strides[1]), dtype=np.uint8),
                      rng.integers(0, 256, (ch, strides[1]), dtype=np.uint8)]
            layers = _layers(rng, w, h)
            host = np.zeros((h, w, 4), np.uint8)
            conv, comp = _jobs(case, [p.ctypes.data for p in planes], host.ctypes.data, layers, [px.ctypes.data for px, *_ in layers],
                               strides)
            assert oracle.oracle_yuv_to_rgb(C.byref(conv)) == 0
            assert oracle.oracle_composite(C.byref(comp)) == 0
            want.append(host)
            dp = [ctx.upload(p) for p in planes]
            dl = [ctx.upload(px) for px, *_ in layers]
            out, out2 = ctx.alloc(w * h * 4), ctx.alloc(w * h * 4)
            keep += dp + dl + [out, out2]
            c, m = _jobs(case, [b.ptr for b in dp], out.ptr, layers, [b.ptr for b in dl], strides)
            conv_d.append(c)
            comp_d.append(m)
            c2, m2 = _jobs(case, [b.ptr for b in dp], out2.ptr, layers, [b.ptr for b in dl], strides)
            conv2_d.append(c2)
            comp2_d.append(m2)
            fused.append(out)
            two_pass.append(out2)
        print(f"yuv + composite: {len(CASES)} jobs, flags {flags}", flush=True)
        ctx.wait(ctx.submit_yuv_composite(conv_d, comp_d, flags))
        ctx.wait(ctx.submit_yuv_to_rgb(conv2_d, flags))
        ctx.wait(ctx.submit_composite(comp2_d, flags))
        for case, out, out2, w_ in zip(CASES, fused, two_pass, want):
            got = ctx.download(out, np.uint8, w_.shape)
            _assert_same(f"{case[:4]} fused", got, w_)
            _assert_same(f"{case[:4]} two-pass", ctx.download(out2, np.uint8, w_.shape), w_)
    finally:
        for b in keep:
            b.free()


def test_yuv_composite_refuses_a_foreign_target(ctx):
    w, h = 64, 16
    planes = [ctx.alloc(w * h), ctx.alloc(w * h // 4), ctx.alloc(w * h // 4)]
    out, other = ctx.alloc(w * h * 4), ctx.alloc(w * h * 4)
    try:
        case = (w, h, w // 2, h // 2, abi.RANGE_LIMITED, abi.MATRIX_BT709)
        conv, comp = _jobs(case, [p.ptr for p in planes], out.ptr, [], [], (w, w // 2))
        comp.target = other.ptr
        with pytest.raises(ReconError) as e:
            ctx.submit_yuv_composite([conv], [comp])
        assert e.value.status == abi.STATUS_INVALID_ARGUMENT
        conv.rgb = out.ptr   # an RGB output as well is refused too
        comp.target = out.ptr
        with pytest.raises(ReconError):
            ctx.submit_yuv_composite([conv], [comp])
    finally:
        for b in planes + [out, other]:
            b.free()
