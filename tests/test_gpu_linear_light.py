"""MMFX linear-light compositing on the device (linear_composite_kernel / linear_convert_kernel)
against the oracle's LinearPixel restatement, bit for bit: several jobs of odd sizes in one
submission and one launch per job, surface and sRGBA8 layers at negative and positive offsets,
clips reaching past the target, opacities 0 / 32768 / 65535 / random, both conversions (including
alpha 0 and channels above alpha, which to_rgba8 clamps), and the reference's group-opacity known
answer (crates/mmfx/src/render.rs:1996-2006) computed on the device."""
import ctypes as C

import numpy as np
import pytest

from mmrecode_amd import abi

pytestmark = pytest.mark.gpu


def _premultiplied(rng, h, w):
    a = rng.integers(0, 65536, (h, w, 1), dtype=np.uint32)
    a[rng.random((h, w, 1)) < 0.1] = 65535
    a[rng.random((h, w, 1)) < 0.05] = 0
    c = (rng.integers(0, 65536, (h, w, 3), dtype=np.uint32) * a // 65535)
    return np.concatenate([c, a], axis=2).astype(np.uint16)


def _layers(rng, W, H):
    out = []
    for k in range(5):
        fmt = abi.LINEAR_SOURCE_SRGBA8 if k % 2 else abi.LINEAR_SOURCE_SURFACE
        w, h = int(rng.integers(1, W + 9)), int(rng.integers(1, H + 9))
        x, y = int(rng.integers(-8, W)), int(rng.integers(-8, H))
        px = (rng.integers(0, 256, (h, w, 4), dtype=np.uint8) if fmt == abi.LINEAR_SOURCE_SRGBA8
              else _premultiplied(rng, h, w))
        clip = (int(rng.integers(-4, W // 2 + 1)), int(rng.integers(-4, H // 2 + 1)), int(rng.integers(W // 2, W + 6)),
                int(rng.integers(H // 2, H + 6)))
        op = [0, 32768, 65535, int(rng.integers(0, 65536)), int(rng.integers(0, 65536))][k]
        out.append((px, fmt, x, y, clip, op))
    return out


def _job(target_ptr, W, H, layers, ptrs):
    j = abi.new(abi.LinearCompositeJob, target=target_ptr, width=W, height=H, stride=W, layer_count=len(layers))
    for i, ((px, fmt, x, y, clip, op), p) in enumerate(zip(layers, ptrs)):
        j.layers[i] = abi.LinearLayer(pixels=p, format=fmt, stride=px.shape[1], x=x, y=y, width=px.shape[1], height=px.shape[0],
                                      clip_left=clip[0], clip_top=clip[1], clip_right=clip[2], clip_bottom=clip[3], opacity=op)
    return j


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_linear_composite_bit_exact(ctx, oracle, flags):
    rng = np.random.default_rng(41 + flags)
    sizes = [(67, 45), (1, 1), (130, 3), (64, 64), (257, 9)]
    jobs, host, dev, keep = [], [], [], []
    for W, H in sizes:
        target = _premultiplied(rng, H, W)
        layers = _layers(rng, W, H)
        want = target.copy()
        hj = _job(want.ctypes.data, W, H, layers, [px.ctypes.data for px, *_ in layers])
        assert oracle.oracle_linear_composite(C.byref(hj)) == 0
        dt = ctx.upload(target)
        dl = [ctx.upload(px) for px, *_ in layers]
        jobs.append(_job(dt.ptr, W, H, layers, [b.ptr for b in dl]))
        host.append(want)
        dev.append(dt)
        keep += dl
    ctx.wait(ctx.submit_linear_composite(jobs, flags))
    for (W, H), want, dt in zip(sizes, host, dev):
        got = ctx.download(dt, np.uint16, want.shape)
        assert np.array_equal(got, want), (W, H, int((got != want).sum()))
    for b in dev + keep:
        b.free()


@pytest.mark.parametrize("direction", [abi.LINEAR_TO_SRGBA8, abi.LINEAR_FROM_SRGBA8])
def test_linear_convert_bit_exact(ctx, oracle, direction):
    rng = np.random.default_rng(7 + direction)
    jobs, checks, bufs = [], [], []
    for W, H in ((67, 45), (1, 1), (300, 7)):
        lin = _premultiplied(rng, H, W)
        if direction == abi.LINEAR_TO_SRGBA8:
            lin[::5, ::3, 0] = 65535          # channels above alpha: to_rgba8 clamps after unpremultiplying
        rgba = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
        hl, hr = lin.copy(), rgba.copy()
        hj = abi.new(abi.LinearConvertJob, direction=direction, width=W, height=H, linear=hl.ctypes.data, linear_stride=W,
                     rgba_stride=W * 4, rgba=hr.ctypes.data)
        assert oracle.oracle_linear_convert(C.byref(hj)) == 0
        dl, dr = ctx.upload(lin), ctx.upload(rgba)
        jobs.append(abi.new(abi.LinearConvertJob, direction=direction, width=W, height=H, linear=dl.ptr, linear_stride=W,
                            rgba_stride=W * 4, rgba=dr.ptr))
        checks.append((dl, dr, hl, hr))
        bufs += [dl, dr]
    ctx.wait(ctx.submit_linear_convert(jobs))
    for dl, dr, hl, hr in checks:
        if direction == abi.LINEAR_TO_SRGBA8:
            assert np.array_equal(ctx.download(dr, np.uint8, hr.shape), hr)
        else:
            assert np.array_equal(ctx.download(dl, np.uint16, hl.shape), hl)
    for b in bufs:
        b.free()


def test_group_opacity_known_answer_on_device(ctx):
    """#000 background, a group holding #f00 blended at opacity 0.5 (32768): [187..188, 0, 0, 255]."""
    target = ctx.upload(np.zeros((1, 1, 4), np.uint16))
    black, red = ctx.upload(np.array([[[0, 0, 0, 255]]], np.uint8)), ctx.upload(np.array([[[255, 0, 0, 255]]], np.uint8))
    group = ctx.upload(np.zeros((1, 1, 4), np.uint16))

    def layer(p, fmt, op):
        return abi.LinearLayer(pixels=p, format=fmt, stride=1, x=0, y=0, width=1, height=1, clip_left=0, clip_top=0,
                               clip_right=1, clip_bottom=1, opacity=op)

    bg = abi.new(abi.LinearCompositeJob, target=target.ptr, width=1, height=1, stride=1, layer_count=1)
    bg.layers[0] = layer(black.ptr, abi.LINEAR_SOURCE_SRGBA8, 65535)
    g = abi.new(abi.LinearCompositeJob, target=group.ptr, width=1, height=1, stride=1, layer_count=1)
    g.layers[0] = layer(red.ptr, abi.LINEAR_SOURCE_SRGBA8, 65535)
    ctx.wait(ctx.submit_linear_composite([bg, g]))
    blend = abi.new(abi.LinearCompositeJob, target=target.ptr, width=1, height=1, stride=1, layer_count=1)
    blend.layers[0] = layer(group.ptr, abi.LINEAR_SOURCE_SURFACE, 32768)
    ctx.wait(ctx.submit_linear_composite([blend]))
    out = ctx.alloc(4)
    ctx.wait(ctx.submit_linear_convert([abi.new(abi.LinearConvertJob, direction=abi.LINEAR_TO_SRGBA8, width=1, height=1,
                                                linear=target.ptr, linear_stride=1, rgba_stride=4, rgba=out.ptr)]))
    px = list(ctx.download(out, np.uint8))
    assert list(ctx.download(target, np.uint16)) == [32768, 0, 0, 65535]
    assert 187 <= px[0] <= 188 and px[1:] == [0, 0, 255], px
    for b in (target, black, red, group, out):
        b.free()
