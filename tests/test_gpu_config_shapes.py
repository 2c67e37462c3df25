"""Bit-exact parity at the benchmark configurations' own shapes:
* MPEG-2: the exact batch bench.py times (1920x1080, 8 closed GOPs x 15 pictures, GOP-interleaved
  decode order, one submission in dependency levels) against the oracle, all 120 pictures;
* H.264: one submission mixing 1280x720, 1920x1088 and 3840x2160 pictures (80x45, 120x68 and
  240x135 macroblocks, in mixed order), weighted bi-prediction and deblocking on, against the
  oracle, with the default deblocking bands and with 4-row bands.
The oracle runs one GOP / picture per thread (its calls release the GIL)."""
import ctypes as C
import threading

import numpy as np
import pytest

from mmrecode_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _threads(jobs):
    ts = [threading.Thread(target=j) for j in jobs]
    for t in ts:
        t.start()
    for t in ts:
        t.join()


def _assert_same(name, got, want):
    if not np.array_equal(got, want):
        diff = np.argwhere(got != want)
        first = tuple(diff[0])
        raise AssertionError(f"{name}: {len(diff)} samples differ; first at {first}: gpu {got[first]} oracle {want[first]}")


def test_mpeg2_bench_batch_bit_exact(ctx, oracle):
    import bench

    w = bench.Mpeg2Workload(ctx)   # 8 GOPs x 15 pictures of 1920x1080, GOP-interleaved
    print(f"mpeg2 bench batch: {w.pictures} pictures", flush=True)
    ctx.wait(w.submit())
    outs = [None] * len(w.gops)

    def run(g):
        o = {}
        for i, u in enumerate(w.gops[g]):
            o[i] = u.host_outputs()
            s = u.host_struct(o[i], wl.mpeg2_ref_ptrs(o, u, "host"))
            assert oracle.oracle_mpeg2_reconstruct(C.byref(s)) == 0
        outs[g] = o

    _threads([lambda g=g: run(g) for g in range(len(w.gops))])
    for (g, i), d in w.dev.items():
        for n in ("y", "cb", "cr"):
            _assert_same(f"GOP {g} picture {i} ({w.gops[g][i].meta['type']}) {n}", d.download(n), outs[g][i][n])
    for d in w.dev.values():
        d.free()


@pytest.mark.parametrize("band_rows", [16, 4])
def test_h264_mixed_sizes_one_submission_bit_exact(ctx, oracle, band_rows):
    sizes = [(240, 135), (80, 45), (120, 68), (80, 45), (240, 135), (120, 68)]
    refs = {}
    for sz in set(sizes):
        host = [dict(zip(("y", "cb", "cr"), wl.h264_reference_planes(2024, sz[0], sz[1], i))) for i in range(4)]
        refs[sz] = (host, [{n: ctx.upload(a) for n, a in r.items()} for r in host])
    units, dev = [], []
    for k, (wmb, hmb) in enumerate(sizes):
        u = wl.h264_picture(2024, 50 + k, wmb, hmb)   # weighted, 40 % bi-predicted, deblocking on
        ptrs = [{n: b.ptr for n, b in r.items()} for r in refs[(wmb, hmb)][1]]
        units.append(u)
        dev.append(wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, ptrs, "dev")))
    saved = ctx.get_tuning()
    ctx.set_tuning(h264_deblock_band_rows=band_rows)
    try:
        print(f"h264 mixed sizes: {len(units)} pictures, {band_rows}-row bands", flush=True)
        ctx.wait(ctx.submit_h264([d.struct for d in dev]))
    finally:
        ctx.set_tuning(**saved)
    want = [None] * len(units)

    def run(k):
        u = units[k]
        o = u.host_outputs()
        s = u.host_struct(o, wl.h264_ref_ptrs(u, refs[(u.meta["width_mbs"], u.meta["height_mbs"])][0], "host"))
        assert oracle.oracle_h264_reconstruct(C.byref(s)) == 0
        want[k] = o

    _threads([lambda k=k: run(k) for k in range(len(units))])
    for k, d in enumerate(dev):
        for n in ("y", "cb", "cr"):
            _assert_same(f"picture {k} {sizes[k]} {n}", d.download(n), want[k][n])
        d.free()
    for _, devr in refs.values():
        for r in devr:
            for b in r.values():
                b.free()


# Output geometries of the mixed workload's frames: (width, height, chroma width, chroma height,
# luma stride, chroma stride) of MJPEG 640x480 4:2:0, MPEG-2 1920x1080 (coded-size planes), DV
# 720x576 4:2:0 and 720x480 4:1:1, H.264 1280x720, 1920x1088 and 3840x2160.
MIXED_GEOMETRIES = [(640, 480, 320, 240, 640, 320), (1920, 1080, 960, 540, 1920, 960), (720, 576, 360, 288, 720, 360),
                    (720, 480, 180, 480, 720, 180), (1280, 720, 640, 360, 1280, 640), (1920, 1088, 960, 544, 1920, 960),
                    (3840, 2160, 1920, 1080, 3840, 1920)]


def test_mixed_workload_convert_and_composite_shapes_bit_exact(ctx, oracle):
    """BT.709 limited-range YUV->RGBA of every output geometry of the mixed workload in one batched
    submission, then a 3-layer premultiplied alpha-over composite onto each converted frame in one
    batched submission, against the oracle's ycbcr_to_rgb and blend chain."""
    from mmrecode_amd import abi

    rng = np.random.default_rng(99)
    host, dev, keep = [], [], []
    conv_d, comp_d = [], []
    for (w, h, cw, ch, ls, cs) in MIXED_GEOMETRIES:
        planes = [rng.integers(0, 256, (h, ls), dtype=np.uint8), rng.integers(0, 256, (ch, cs), dtype=np.uint8),
                  rng.integers(0, 256, (ch, cs), dtype=np.uint8)]
        layers = []
        for _ in range(3):
            px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
            px[..., :3] = (px[..., :3].astype(np.uint32) * px[..., 3:4] // 255).astype(np.uint8)
            layers.append(px)
        rgba = np.zeros((h, w, 4), np.uint8)
        job = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=ls,
                      chroma_stride=cs, range=abi.RANGE_LIMITED, matrix=abi.MATRIX_BT709, rgba=rgba.ctypes.data)
        for k in range(3):
            job.planes[k] = planes[k].ctypes.data
        assert oracle.oracle_yuv_to_rgb(C.byref(job)) == 0
        cj = abi.new(abi.CompositeJob, target=rgba.ctypes.data, width=w, height=h, stride=w * 4, layer_count=3)
        for i, px in enumerate(layers):
            cj.layers[i] = abi.Layer(rgba=px.ctypes.data, width=w, height=h, stride=w * 4, x=0, y=0,
                                     op=abi.BLEND_PREMULTIPLIED_OVER, opacity=255)
        assert oracle.oracle_composite(C.byref(cj)) == 0
        host.append(rgba)
        bufs = [ctx.upload(p) for p in planes]
        lb = [ctx.upload(px) for px in layers]
        out = ctx.alloc(w * h * 4)
        keep += bufs + lb + [out]
        dj = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=ls,
                     chroma_stride=cs, range=abi.RANGE_LIMITED, matrix=abi.MATRIX_BT709, rgba=out.ptr)
        for k in range(3):
            dj.planes[k] = bufs[k].ptr
        conv_d.append(dj)
        dc = abi.new(abi.CompositeJob, target=out.ptr, width=w, height=h, stride=w * 4, layer_count=3)
        for i, b in enumerate(lb):
            dc.layers[i] = abi.Layer(rgba=b.ptr, width=w, height=h, stride=w * 4, x=0, y=0,
                                     op=abi.BLEND_PREMULTIPLIED_OVER, opacity=255)
        comp_d.append(dc)
        dev.append(out)
    print(f"mixed geometries: {len(dev)} frames", flush=True)
    ctx.wait(ctx.submit_yuv_to_rgb(conv_d))
    ctx.wait(ctx.submit_composite(comp_d))
    for g, out, want in zip(MIXED_GEOMETRIES, dev, host):
        _assert_same(f"{g[0]}x{g[1]}", ctx.download(out, np.uint8, want.shape), want)
    for b in keep:
        b.free()


def test_mixed_workload_step_bit_exact(ctx, oracle):
    """One step of bench.py's mixed workload (30 MJPEG, a 15-picture MPEG-2 GOP, 30 DV, 12 H.264
    pictures up to 3840x2160; one batch per codec, then conversion and composite batches, the
    3840x2160 H.264 chain and the others on two streams) against the oracle chain of every frame, RGBA bit for bit."""
    import bench

    w = bench.MixedWorkload(ctx)
    print(f"mixed step: {w.pictures} frames", flush=True)
    w.wait(w.submit())   # the H.264 chain on ctx's stream, the others on a second context's
    outs = {i: u.host_outputs() for i, (_, u, _) in enumerate(w.frames)}
    want = [None] * w.pictures
    mpeg2 = [i for i, (c, _, _) in enumerate(w.frames) if c == bench.abi.CODEC_MPEG2]
    rest = [i for i in range(w.pictures) if i not in mpeg2]

    def run(idx):
        host = {}
        for i in idx:
            want[i] = w.oracle_chain(oracle, i, outs[i], host)

    _threads([lambda: run(mpeg2)] + [lambda k=k: run(rest[k::15]) for k in range(15)])
    for i in range(w.pictures):
        got = ctx.download(w.rgba[i], np.uint8, want[i].shape)
        _assert_same(f"frame {i} (codec {w.frames[i][0]})", got, want[i])
    w.close()
