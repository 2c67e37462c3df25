"""Benchmark of the MI355X reconstruction path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mpeg2|mjpeg|dv|h264|h264i]

A step is one pass of the hot path over one batch of synthetic post-entropy input, all inputs
resident in HBM before the timed region:
  mpeg2 (default)  8 closed GOPs x 15 pictures of 1920x1080 4:2:0 in one persistent launch
  mjpeg            120 frames of 640x480 4:2:0 + RGB8 in one launch
  dv               256 frames, 720x576 4:2:0 and 720x480 4:1:1 interleaved, one launch
  h264             16 pictures each of 1280x720, 1920x1080, 3840x2160 with deblocking, one launch
  h264i            the same pictures with every macroblock intra (I pictures)
Reported beside the batched value: the plain path (one picture per launch, synchronous) and the
CPU baseline (the oracle, a scalar C restatement of the reference functions, on the host cores
over a bounded sample of the same workload).
"""
import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import mmrecode_amd  # noqa: E402
from mmrecode_amd import abi, metrics  # noqa: E402
from mmrecode_amd import workloads as wl  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per step of the dominant kernel, measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and
# scaled by the counters' reading for a device-to-device copy of known size (see the summaries in
# profiles/); absent entries report traffic as null.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
FP64_PEAK_TFLOPS = 78.6


def cpu_passes(jobs, min_cpu_seconds=12.0, min_seconds=1.0, max_passes=2000):
    """Runs `jobs` (callables, one thread each; the oracle calls release the GIL) as passes until
    at least `min_cpu_seconds` of thread time (wall time x threads) and `min_seconds` of wall time
    have passed. Returns (passes, seconds)."""
    passes, t = 0, time.time()
    while True:
        ts = [threading.Thread(target=j) for j in jobs]
        for th in ts:
            th.start()
        for th in ts:
            th.join()
        passes += 1
        dt = time.time() - t
        if (dt * len(jobs) >= min_cpu_seconds and dt >= min_seconds) or passes >= max_passes:
            return passes, dt


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """(threads this process may run on, os.cpu_count(), CPU model string)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return usable, os.cpu_count(), model


def percentile(values, q):
    return float(np.percentile(np.asarray(values, float), q)) if len(values) else None


# ------------------------------------------------------------------ workloads
class Mpeg2Workload:
    name = "mpeg2"
    codec = abi.CODEC_MPEG2
    metric_unit = "Mpix/s"
    kernel = "mpeg2_kernel"
    phases = ("mpeg2_order_kernel", "mpeg2_kernel")

    def __init__(self, ctx, gops=8, width=1920, height=1080, seed=2024):
        self.ctx, self.width, self.height = ctx, width, height
        t = time.time()
        self.gops = [wl.mpeg2_gop(seed, g, width, height) for g in range(gops)]
        log(f"generated {gops} GOPs in {time.time() - t:.1f}s")
        # GOP-interleaved decode order: picture i of every GOP, then i + 1, ... (anchor chains of
        # different GOPs run side by side).
        self.order = [(g, i) for i in range(len(self.gops[0])) for g in range(gops)]
        self.dev = {}
        self.structs = []
        pos = {}
        t = time.time()
        for k, (g, i) in enumerate(self.order):
            u = self.gops[g][i]
            extra = {"forward_source": -1, "backward_source": -1}
            for tag, src in (("fwd", "forward_source"), ("bwd", "backward_source")):
                ref = u.meta[tag]
                if ref is not None:
                    extra[src] = pos[(g, ref)]
                    for n in ("y", "cb", "cr"):
                        extra[f"{tag}_{n}"] = self.dev[(g, ref)].ptr(n)
            d = wl.DeviceUnit(ctx, u, extra)
            self.dev[(g, i)] = d
            pos[(g, i)] = k
            self.structs.append(d.struct)
        log(f"uploaded {len(self.structs)} pictures in {time.time() - t:.1f}s")
        self.pictures = len(self.structs)
        self.mpix = self.pictures * metrics.mpix(width, height)
        self.bytes = sum(metrics.mpeg2_picture_bytes(self.gops[g][i]) for g, i in self.order)
        self.kernel_bytes = {"mpeg2_kernel": self.bytes}   # the order kernel re-reads records only
        # f64 IDCT operation count at the reference's order (dense: 1024 mul + 1024 add per block)
        self.f64_ops = sum(self.gops[g][i].meta["coded_blocks"] for g, i in self.order) * 2048

    def submit(self, flags=0):
        return self.ctx.submit_mpeg2(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_mpeg2(items, flags)

    def config(self):
        return {"workload": f"MPEG-2 MP@HL {self.width}x{self.height} 4:2:0, {len(self.gops)} closed GOPs x 15 "
                            f"pictures per launch", "pictures_per_step": self.pictures,
                "resolution": f"{self.width}x{self.height}", "gop": wl.GOP_DECODE_ORDER}

    def cpu_sample(self, threads):
        """One GOP per thread through the oracle, Mpix/s: the step's GOPs plus further GOPs of the
        same configuration (seeded after them) when there are more threads than GOPs."""
        from oracle import binding

        L = binding.lib()
        n = max(1, threads)
        gops = self.gops + [wl.mpeg2_gop(2024, g, self.width, self.height) for g in range(len(self.gops), n)]

        outs = [{i: u.host_outputs() for i, u in enumerate(gops[g])} for g in range(n)]

        def run_gop(g):
            for i, u in enumerate(gops[g]):
                s = u.host_struct(outs[g][i], wl.mpeg2_ref_ptrs(outs[g], u, "host"))
                assert L.oracle_mpeg2_reconstruct(C.byref(s)) == 0

        passes, dt = cpu_passes([lambda g=g: run_gop(g) for g in range(n)])
        mp = passes * sum(len(gops[g]) for g in range(n)) * metrics.mpix(self.width, self.height)
        return mp / dt, n, (f"{passes} pass(es) over {n} GOP(s) x 15 pictures {self.width}x{self.height}, one GOP per "
                            f"thread, {dt:.1f} s wall on {n} threads")


class MjpegWorkload:
    name = "mjpeg"
    codec = abi.CODEC_MJPEG
    metric_unit = "Mpix/s"
    kernel = "mjpeg_blocks_kernel"
    phases = ("mjpeg_blocks_kernel", "yuv_to_rgb_kernel")

    def __init__(self, ctx, frames=120, seed=2024):
        self.ctx = ctx
        self.units = [wl.mjpeg_frame(seed, i) for i in range(frames)]
        self.dev = [wl.DeviceUnit(ctx, u) for u in self.units]
        self.structs = [d.struct for d in self.dev]
        self.pictures = frames
        self.mpix = frames * metrics.mpix(640, 480)
        self.bytes = sum(metrics.mjpeg_frame_bytes(u) for u in self.units)
        self.kernel_bytes = {"mjpeg_blocks_kernel": sum(metrics.mjpeg_block_kernel_bytes(u) for u in self.units),
                             "yuv_to_rgb_kernel": sum(metrics.mjpeg_rgb_kernel_bytes(u) for u in self.units)}
        self.f64_ops = sum(u.meta["blocks"] for u in self.units) * 2048

    def submit(self, flags=0):
        return self.ctx.submit_mjpeg(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_mjpeg(items, flags)

    def config(self):
        return {"workload": "MJPEG baseline 640x480 4:2:0 + BT.601 RGB8, 120 frames per launch",
                "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        n = max(1, min(threads, len(self.units)))
        outs = [u.host_outputs() for u in self.units[:n]]
        passes, dt = cpu_passes([lambda u=u, o=o: L.oracle_mjpeg_reconstruct(C.byref(u.host_struct(o)))
                                 for u, o in zip(self.units[:n], outs)])
        return (passes * n * metrics.mpix(640, 480) / dt, n,
                f"{passes} pass(es) over {n} frames 640x480 (+RGB), one frame per thread, {dt:.1f} s wall on {n} threads")


class DvWorkload:
    name = "dv"
    codec = abi.CODEC_DV
    metric_unit = "Mpix/s"
    kernel = "dv_kernel"

    def __init__(self, ctx, frames=256, seed=2024):
        self.ctx = ctx
        rng = np.random.default_rng(seed)
        systems = rng.integers(0, 2, frames)
        self.units = [wl.dv_frame(seed, i, system=int(s)) for i, s in enumerate(systems)]
        self.dev = [wl.DeviceUnit(ctx, u) for u in self.units]
        self.structs = [d.struct for d in self.dev]
        self.pictures = frames
        self.mpix = sum(metrics.mpix(720, u.meta["height"]) for u in self.units)
        self.bytes = sum(metrics.dv_frame_bytes(u) for u in self.units)
        self.f64_ops = None

    def submit(self, flags=0):
        return self.ctx.submit_dv(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_dv(items, flags)

    def config(self):
        return {"workload": "DV25 720x576 4:2:0 + 720x480 4:1:1 interleaved, 256 frames per launch",
                "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        n = max(1, min(threads, len(self.units)))
        outs = [u.host_outputs() for u in self.units[:n]]
        passes, dt = cpu_passes([lambda u=u, o=o: L.oracle_dv_reconstruct(C.byref(u.host_struct(o)))
                                 for u, o in zip(self.units[:n], outs)])
        mp = passes * sum(metrics.mpix(720, u.meta["height"]) for u in self.units[:n])
        return mp / dt, n, f"{passes} pass(es) over {n} DV frames, one frame per thread, {dt:.1f} s wall on {n} threads"


class H264Workload:
    name = "h264"
    codec = abi.CODEC_H264
    metric_unit = "Mpix/s"
    kernel = "h264_mb_kernel"
    SIZES = ((80, 45), (120, 68), (240, 135))

    def phase_names(self, n):
        if n == 4:
            return ("h264_bs_kernel", "h264_mb_kernel", "h264_intra_kernel", "h264_deblock_kernel")
        return ("h264_bs_kernel", "h264_mb_kernel", "h264_deblock_kernel")[:n]

    picture_kwargs = {}

    def __init__(self, ctx, per_size=16, seed=2024):
        self.ctx = ctx
        self.host_refs, self.dev_refs = {}, {}
        self.units, self.structs = [], []
        self.dev = []
        t = time.time()
        for wmb, hmb in self.SIZES:
            host = [dict(zip(("y", "cb", "cr"), wl.h264_reference_planes(seed, wmb, hmb, i))) for i in range(4)]
            devr = [{n: ctx.upload(a) for n, a in r.items()} for r in host]
            self.host_refs[(wmb, hmb)] = host
            self.dev_refs[(wmb, hmb)] = devr
            ptrs = [{n: b.ptr for n, b in r.items()} for r in devr]
            for i in range(per_size):
                u = wl.h264_picture(seed, i, wmb, hmb, **self.picture_kwargs)
                d = wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, ptrs, "dev"))
                self.units.append(u)
                self.dev.append(d)
                self.structs.append(d.struct)
        log(f"generated and uploaded {len(self.units)} H.264 pictures in {time.time() - t:.1f}s")
        self.pictures = len(self.units)
        self.mpix = sum(metrics.mpix(u.meta["width_mbs"] * 16, u.meta["height_mbs"] * 16) for u in self.units)
        self.bytes = sum(metrics.h264_picture_bytes(u) for u in self.units)
        self.f64_ops = None

    def submit(self, flags=0):
        return self.ctx.submit_h264(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_h264(items, flags)

    def config(self):
        return {"workload": "H.264 High 4:2:0 inter + deblocking, 16 pictures each of 1280x720, 1920x1088, "
                            "3840x2160 per launch", "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        picks = [u for u in self.units if u.meta["width_mbs"] == 120][: max(1, min(threads, 16))]
        outs = [u.host_outputs() for u in picks]

        def run(u, o):
            s = u.host_struct(o, wl.h264_ref_ptrs(u, self.host_refs[(u.meta["width_mbs"], u.meta["height_mbs"])], "host"))
            assert L.oracle_h264_reconstruct(C.byref(s)) == 0

        passes, dt = cpu_passes([lambda u=u, o=o: run(u, o) for u, o in zip(picks, outs)])
        mp = passes * len(picks) * metrics.mpix(1920, 1088)
        return mp / dt, len(picks), (f"{passes} pass(es) over {len(picks)} 1920x1088 pictures, one picture per thread, "
                                     f"{dt:.1f} s wall on {len(picks)} threads")


class H264IntraWorkload(H264Workload):
    """The H.264 sizes with every macroblock intra (I pictures: Intra4x4 / Intra8x8 / Intra16x16 /
    I_PCM mixed 40 / 25 / 25 / 10 %, one slice, deblocking on): the intra pass's wavefront."""
    name = "h264i"
    kernel = "h264_intra_kernel"
    picture_kwargs = {"intra_fraction": 1.0}

    def config(self):
        return {"workload": "H.264 High 4:2:0 intra (I pictures) + deblocking, 16 pictures each of 1280x720, "
                            "1920x1088, 3840x2160 per launch", "pictures_per_step": self.pictures}


# ------------------------------------------------------------------ compositor neighbours
class ScaleWorkload:
    """The canvas scaler (scale_yuv420_to_canvas_cpu, crates/render/src/compositor.rs:89-141) with
    Lanczos3: 16 frames 1920x1080 -> 1280x720 (FIT, the preview downscale) and 16 frames 1280x720
    -> 1920x1080 (FIT, the export upscale) per launch. Mpix counts canvas samples (luma)."""
    name = "scale"
    codec = None
    data = "synthetic (seeded random planes)"
    metric = "scaled canvas Mpix/s (4:2:0 Lanczos3, centred crop / pad)"
    metric_unit = "Mpix/s"
    kernel = "scale_output_kernel"
    dtype = "u8 in/out, f32 (Lanczos3 taps and sums) arithmetic"
    CASES = ((1920, 1080, 1280, 720), (1280, 720, 1920, 1080))

    def __init__(self, ctx, per_case=16, seed=2024):
        self.ctx = ctx
        rng = np.random.default_rng(seed)
        self.host, self.bufs, self.structs = [], [], []
        for w, h, cw, ch in self.CASES:
            for _ in range(per_case):
                planes = [rng.integers(0, 256, s, dtype=np.uint8) for s in ((h, w), (h // 2, w // 2), (h // 2, w // 2))]
                src = [ctx.upload(p) for p in planes]
                dst = [ctx.alloc(cw * ch), ctx.alloc(cw * ch // 4), ctx.alloc(cw * ch // 4)]
                j = abi.new(abi.ScaleJob, width=w, height=h, canvas_width=cw, canvas_height=ch, mode=abi.SCALE_FIT,
                            range=abi.RANGE_LIMITED, filter=abi.SCALE_FILTER_LANCZOS3)
                for k in range(3):
                    j.src[k], j.src_stride[k] = src[k].ptr, planes[k].shape[1]
                    j.dst[k], j.dst_stride[k] = dst[k].ptr, cw >> (1 if k else 0)
                self.host.append((planes, w, h, cw, ch))
                self.bufs += src + dst
                self.structs.append(j)
        self.pictures = len(self.structs)
        self.mpix = sum(metrics.mpix(cw, ch) for _, _, _, cw, ch in self.host)
        self.bytes = sum(1.5 * (w * h + cw * ch) for _, w, h, cw, ch in self.host)
        self.f64_ops = None

    def submit(self, flags=0):
        return self.ctx.submit_scale(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_scale(items, flags)

    def config(self):
        return {"workload": "4:2:0 canvas scaling, Lanczos3 FIT: 16 frames 1920x1080 -> 1280x720 and 16 frames "
                            "1280x720 -> 1920x1080 per launch", "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        picks = [self.host[0], self.host[-1]] * max(1, threads // 2)
        picks = picks[:max(1, threads)]
        outs = []
        for planes, w, h, cw, ch in picks:
            out = [np.zeros((ch, cw), np.uint8), np.zeros((ch // 2, cw // 2), np.uint8), np.zeros((ch // 2, cw // 2), np.uint8)]
            j = abi.new(abi.ScaleJob, width=w, height=h, canvas_width=cw, canvas_height=ch, mode=abi.SCALE_FIT,
                        range=abi.RANGE_LIMITED, filter=abi.SCALE_FILTER_LANCZOS3)
            for k in range(3):
                j.src[k], j.src_stride[k] = planes[k].ctypes.data, planes[k].shape[1]
                j.dst[k], j.dst_stride[k] = out[k].ctypes.data, out[k].shape[1]
            outs.append((j, out))
        passes, dt = cpu_passes([lambda j=j: L.oracle_scale_yuv420(C.byref(j)) for j, _ in outs])
        mp = passes * sum(metrics.mpix(cw, ch) for _, _, _, cw, ch in picks)
        return mp / dt, len(picks), (f"{passes} pass(es) over {len(picks)} frames (half 1080p->720p, half 720p->1080p), "
                                     f"one frame per thread, {dt:.1f} s wall on {len(picks)} threads")


class OverlayWorkload:
    """Overlay preparation (PreparedOverlay::prepare_yuv, crates/render/src/project_compositor.rs:1098-1142):
    32 RGBA overlays of 1920x1080 per launch into the seven prepared planes."""
    name = "overlay"
    codec = None
    data = "synthetic (seeded random planes)"
    metric = "prepared overlay Mpix/s (RGBA -> BT.709 limited + full Y, alpha-weighted U/V, chroma alpha)"
    metric_unit = "Mpix/s"
    kernel = "overlay_prepare_kernel"
    dtype = "u8 in/out, int32 arithmetic"

    def __init__(self, ctx, frames=32, width=1920, height=1080, seed=2024):
        self.ctx = ctx
        rng = np.random.default_rng(seed)
        self.host, self.bufs, self.structs = [], [], []
        cw, ch = (width + 1) // 2, (height + 1) // 2
        for _ in range(frames):
            rgba = rng.integers(0, 256, (height, width, 4), dtype=np.uint8)
            d = ctx.upload(rgba)
            outs = {n: ctx.alloc(width * height if n.startswith("y_") else cw * ch)
                    for n in ("y_limited", "y_full", "u_limited", "v_limited", "u_full", "v_full", "chroma_alpha")}
            self.structs.append(abi.new(abi.OverlayPrepareJob, rgba=d.ptr, width=width, height=height,
                                        **{n: b.ptr for n, b in outs.items()}))
            self.host.append(rgba)
            self.bufs += [d] + list(outs.values())
        self.pictures = frames
        self.mpix = frames * metrics.mpix(width, height)
        self.bytes = frames * (width * height * (4 + 2) + cw * ch * 5)
        self.f64_ops = None

    def submit(self, flags=0):
        return self.ctx.submit_overlay_prepare(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_overlay_prepare(items, flags)

    def config(self):
        return {"workload": "overlay preparation, 32 RGBA overlays of 1920x1080 per launch", "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        n = max(1, min(threads, len(self.host)))
        jobs = []
        for rgba in self.host[:n]:
            h, w = rgba.shape[:2]
            outs = {k: np.zeros(w * h if k.startswith("y_") else ((w + 1) // 2) * ((h + 1) // 2), np.uint8)
                    for k in ("y_limited", "y_full", "u_limited", "v_limited", "u_full", "v_full", "chroma_alpha")}
            j = abi.new(abi.OverlayPrepareJob, rgba=rgba.ctypes.data, width=w, height=h, **{k: v.ctypes.data for k, v in outs.items()})
            jobs.append((j, outs))
        passes, dt = cpu_passes([lambda j=j: L.oracle_overlay_prepare(C.byref(j)) for j, _ in jobs])
        mp = passes * n * metrics.mpix(self.host[0].shape[1], self.host[0].shape[0])
        return mp / dt, n, f"{passes} pass(es) over {n} 1920x1080 overlays, one per thread, {dt:.1f} s wall on {n} threads"


# ------------------------------------------------------------------ served mixed workload
class ServedMix:
    """Open-loop served workload: Poisson request arrivals over 64 streams (16 per codec), each
    request a window of 1-30 consecutive frames of its stream, every frame followed by BT.709
    limited-range YUV->RGBA and a 3-layer premultiplied alpha-over composite at the output size.
    Inputs are pooled per codec and resident in HBM; every stream owns a ring of 32 output slots, so
    MPEG-2 streams keep their anchors resident across requests, and H.264 streams reference 4
    resident pictures per resolution."""

    RING = 32
    MAX_WINDOW = 30
    H264_SIZES = ((80, 45), (120, 68), (240, 135))

    def __init__(self, ctx, seed=2024, streams=64, hbm_budget=96 << 30):
        self.ctx, self.seed, self.hbm_budget = ctx, seed, hbm_budget
        t = time.time()
        self.pool = {abi.CODEC_MJPEG: [wl.mjpeg_frame(seed, 1000 + i, rgb=False) for i in range(30)],
                     abi.CODEC_DV: [wl.dv_frame(seed, 1000 + i, system=i % 2) for i in range(30)],
                     abi.CODEC_MPEG2: wl.mpeg2_gop(seed, 99)}
        self.h264_pool = {sz: [wl.h264_picture(seed, 100 + i, *sz) for i in range(4)] for sz in self.H264_SIZES}
        self.h264_refs = {}
        for sz in self.H264_SIZES:
            refs = [dict(zip(("y", "cb", "cr"), wl.h264_reference_planes(seed, sz[0], sz[1], i))) for i in range(4)]
            self.h264_refs[sz] = [{n: ctx.upload(a).ptr for n, a in r.items()} for r in refs]
        # pooled inputs, uploaded once: id(unit) -> name -> device pointer
        self.inputs = {}
        for units in list(self.pool.values()) + list(self.h264_pool.values()):
            for u in units:
                self.inputs[id(u)] = {k: ctx.upload(v).ptr for k, v in u.inputs.items()}
        self.layers = {}
        self.streams = []
        for sid in range(streams):
            codec = (abi.CODEC_MJPEG, abi.CODEC_MPEG2, abi.CODEC_DV, abi.CODEC_H264)[sid % 4]
            size = self.H264_SIZES[(sid // 4) % 3] if codec == abi.CODEC_H264 else None
            self.streams.append({"id": sid + 1, "codec": codec, "size": size, "pos": 0, "ring": {}})
        log(f"served mix: pools generated and uploaded in {time.time() - t:.1f}s")

    # ---- geometry
    def _units_of(self, st):
        return self.h264_pool[st["size"]] if st["codec"] == abi.CODEC_H264 else self.pool[st["codec"]]

    @staticmethod
    def geometry(codec, u):
        """(width, height, chroma_w, chroma_h, luma_stride, chroma_stride, plane names) of the output."""
        if codec == abi.CODEC_MJPEG:
            w, h = u.meta["width"], u.meta["height"]
            return w, h, (w + 1) // 2, (h + 1) // 2, w, (w + 1) // 2, ("plane0", "plane1", "plane2")
        if codec == abi.CODEC_DV:
            h = u.meta["height"]
            cw, ch = (360, 288) if u.meta["system"] == abi.DV_SYSTEM_625_50 else (180, 480)
            return 720, h, cw, ch, 720, cw, ("y", "cb", "cr")
        if codec == abi.CODEC_MPEG2:
            w, h = u.meta["width"], u.meta["height"]
            W = -(-w // 16) * 16
            return w, h, (w + 1) // 2, (h + 1) // 2, W, W // 2, ("y", "cb", "cr")
        W, H = u.meta["width_mbs"] * 16, u.meta["height_mbs"] * 16
        return W, H, W // 2, H // 2, W, W // 2, ("y", "cb", "cr")

    def _layers(self, w, h):
        if (w, h) not in self.layers:
            rng = np.random.Generator(np.random.PCG64([self.seed, 77, w, h]))
            bufs = []
            for _ in range(3):
                px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
                px[..., :3] = (px[..., :3].astype(np.uint32) * px[..., 3:4] // 255).astype(np.uint8)
                bufs.append(self.ctx.upload(px))
            self.layers[(w, h)] = bufs
        return self.layers[(w, h)]

    def _slot(self, st, u, k):
        """Output buffers of ring slot k of stream st, allocated on first use at the largest size any
        frame of the stream's pool needs (DV streams mix 720x576 4:2:0 and 720x480 4:1:1 frames)."""
        if k not in st["ring"]:
            units = self._units_of(st)
            need = {}
            for v in units:
                for n, (shape, dt, _) in v.outputs.items():
                    need[n] = max(need.get(n, 0), int(np.prod(shape)) * np.dtype(dt).itemsize)
                w, h = self.geometry(st["codec"], v)[:2]
                need["rgba"] = max(need.get("rgba", 0), w * h * 4)
            st["ring"][k] = {n: self.ctx.alloc(b) for n, b in need.items()}
        slot = st["ring"][k]
        for n, (shape, dt, _) in u.outputs.items():
            assert slot[n].nbytes >= int(np.prod(shape)) * np.dtype(dt).itemsize, (n, shape)
        w, h = self.geometry(st["codec"], u)[:2]
        assert slot["rgba"].nbytes >= w * h * 4
        return slot

    def frame(self, st, p, start):
        """Descriptor, convert job, composite job, bytes and Mpix of picture p of stream st; `start`
        is the stream position of the request's first frame."""
        codec, units = st["codec"], self._units_of(st)
        u = units[p % len(units)]
        slot = self._slot(st, u, p % self.RING)
        ptrs = dict(self.inputs[id(u)])
        ptrs.update({n: b.ptr for n, b in slot.items() if n != "rgba"})
        if codec == abi.CODEC_MPEG2:
            ptrs.update({"forward_source": -1, "backward_source": -1})
            gop0 = p - p % len(units)
            for tag, src in (("fwd", "forward_source"), ("bwd", "backward_source")):
                ref = u.meta[tag]
                if ref is None:
                    continue
                rp = gop0 + ref
                if rp >= start:
                    ptrs[src] = rp - start
                rs = self._slot(st, units[ref], rp % self.RING)
                for n in ("y", "cb", "cr"):
                    ptrs[f"{tag}_{n}"] = rs[n].ptr
        elif codec == abi.CODEC_H264:
            ptrs.update(wl.h264_ref_ptrs(u, self.h264_refs[st["size"]], "dev"))
        desc = u.fill(ptrs)
        w, h, cw, ch, ls, cs, names = self.geometry(codec, u)
        conv = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=ls,
                       chroma_stride=cs, range=abi.RANGE_LIMITED, matrix=abi.MATRIX_BT709, rgba=slot["rgba"].ptr)
        for k, n in enumerate(names):
            conv.planes[k] = slot[n].ptr
        comp = abi.new(abi.CompositeJob, target=slot["rgba"].ptr, width=w, height=h, stride=w * 4, layer_count=3)
        for i, b in enumerate(self._layers(w, h)):
            comp.layers[i] = abi.Layer(rgba=b.ptr, width=w, height=h, stride=w * 4, x=0, y=0,
                                       op=abi.BLEND_PREMULTIPLIED_OVER, opacity=255)
        nbytes = u.input_bytes() + u.output_bytes() + w * h * 4
        return desc, conv, comp, nbytes, metrics.mpix(w, h)

    def trace(self, rate, seconds, rng):
        """Arrival times and requests: (t, stream, frames)."""
        out, t = [], 0.0
        by_codec = {c: [s for s in self.streams if s["codec"] == c] for c in Server_CODECS}
        while True:
            t += rng.exponential(1.0 / rate)
            if t >= seconds:
                break
            codec = Server_CODECS[rng.integers(0, 4)]
            st = by_codec[codec][rng.integers(0, len(by_codec[codec]))]
            out.append((t, st, int(rng.integers(1, self.MAX_WINDOW + 1))))
        return out

    def build(self, trace):
        """Descriptors of every request of a trace (stream positions advance in arrival order)."""
        reqs = []
        for t, st, n in trace:
            start = st["pos"]
            frames = [self.frame(st, start + i, start) for i in range(n)]
            st["pos"] += n
            struct = Server.STRUCTS[st["codec"]]
            farr = (struct * n)(*[f[0] for f in frames])
            carr = (abi.YuvToRgbJob * n)(*[f[1] for f in frames])
            parr = (abi.CompositeJob * n)(*[f[2] for f in frames])
            req = abi.new(abi.Request, codec=st["codec"], stream_id=st["id"], frame_count=n,
                          frames=C.addressof(farr), convert=C.addressof(carr), composite=C.addressof(parr),
                          hbm_bytes=sum(f[3] for f in frames), user_tag=len(reqs))
            reqs.append({"t": t, "req": req, "keep": (farr, carr, parr), "codec": st["codec"], "n": n,
                         "mpix": [f[4] for f in frames], "frames": farr, "convert": carr, "composite": parr})
        return reqs

    def run_served(self, reqs, slots=4):
        ctx, lib = self.ctx, self.ctx.lib
        server = mmrecode_amd.Server(ctx, hbm_budget_bytes=self.hbm_budget, slots=slots, max_batch_frames=64,
                                     max_frames_per_request_per_batch=8)
        lat, mpix, status_bad = [], 0.0, 0
        buf = (abi.FrameCompletion * 8192)()
        n_out = C.c_uint32()

        def poll():
            nonlocal mpix, status_bad
            mmrecode_amd.check(lib.mmr_recon_server_poll(server.handle, buf, 8192, C.byref(n_out)))
            for c in buf[: n_out.value]:
                if c.status != abi.STATUS_OK:
                    status_bad += 1
                lat.append(c.latency_ms)
                mpix += reqs[c.user_tag]["mpix"][c.frame_index]
            return n_out.value

        rid = C.c_uint64()
        t0 = time.perf_counter()
        try:
            for r in reqs:
                while True:
                    now = time.perf_counter() - t0
                    if now >= r["t"]:
                        break
                    if not poll():
                        time.sleep(min(2e-4, r["t"] - now))
                mmrecode_amd.check(lib.mmr_recon_server_enqueue(server.handle, C.byref(r["req"]), C.byref(rid)))
            server.drain()
            while poll():
                pass
            elapsed = time.perf_counter() - t0
            stats = server.stats()
        finally:
            server.close()
        return {"mpix_per_s": round(mpix / elapsed, 1), "frames": len(lat), "bad_status": status_bad,
                "p50_frame_ms": round(percentile(lat, 50), 3), "p99_frame_ms": round(percentile(lat, 99), 3),
                "wall_s": round(elapsed, 3), "batches": stats["batches_launched"],
                "peak_admitted_gb": round(stats["peak_admitted_bytes"] / 1e9, 2)}

    def run_plain(self, reqs, limit_s):
        """The same trace through the plain path: one synchronous launch per frame (decode, then
        convert, then composite), requests in arrival order; stops after limit_s of wall time."""
        ctx = self.ctx
        submit = {abi.CODEC_MJPEG: ctx.submit_mjpeg, abi.CODEC_MPEG2: ctx.submit_mpeg2, abi.CODEC_DV: ctx.submit_dv,
                  abi.CODEC_H264: ctx.submit_h264}
        lat, mpix, done_frames, total = [], 0.0, 0, sum(r["n"] for r in reqs)
        flags = abi.SUBMIT_PER_FRAME_LAUNCH
        t0 = time.perf_counter()
        for r in reqs:
            now = time.perf_counter() - t0
            if now > limit_s:
                break
            if now < r["t"]:
                time.sleep(r["t"] - now)
            for i in range(r["n"]):
                ctx.wait(submit[r["codec"]]([r["frames"][i]], flags))
                ctx.wait(ctx.submit_yuv_to_rgb([r["convert"][i]], flags))
                ctx.wait(ctx.submit_composite([r["composite"][i]], flags))
                lat.append((time.perf_counter() - t0 - r["t"]) * 1e3)
                mpix += r["mpix"][i]
                done_frames += 1
        elapsed = time.perf_counter() - t0
        return {"mpix_per_s": round(mpix / elapsed, 1), "frames": done_frames, "frames_offered": total,
                "p50_frame_ms": round(percentile(lat, 50), 3) if lat else None,
                "p99_frame_ms": round(percentile(lat, 99), 3) if lat else None, "wall_s": round(elapsed, 3)}


Server_CODECS = (abi.CODEC_MJPEG, abi.CODEC_MPEG2, abi.CODEC_DV, abi.CODEC_H264)
Server = mmrecode_amd.Server


def run_served_sweep(ctx, rates, seconds, plain_limit):
    mix = ServedMix(ctx)
    rng = np.random.default_rng(4242)
    out = []
    for rate in rates:
        reqs = mix.build(mix.trace(rate, seconds, rng))
        served = mix.run_served(reqs)
        log(f"served {rate} req/s: {served}")
        entry = {"rate_req_s": rate, "requests": len(reqs), "served": served}
        if plain_limit > 0:
            entry["plain"] = mix.run_plain(reqs, plain_limit)
            log(f"plain  {rate} req/s: {entry['plain']}")
            if entry["plain"]["mpix_per_s"]:
                entry["served_vs_plain"] = round(served["mpix_per_s"] / entry["plain"]["mpix_per_s"], 2)
        out.append(entry)
    return {"config": "Poisson arrivals, 64 streams (16 per codec: MJPEG 640x480, MPEG-2 1920x1080 GOP-15, "
                      "DV 720x576/720x480, H.264 1280x720/1920x1088/3840x2160), windows of 1-30 frames, each frame "
                      "+ BT.709 limited YUV->RGBA + 3-layer premultiplied alpha-over, 96 GB admission budget, "
                      "resident references",
            "seconds_per_rate": seconds, "rates": out}


class MixedWorkload:
    """The served mix's per-frame chain, batched on two streams: a fixed set of request windows -- 30
    MJPEG 640x480 frames, one MPEG-2 1920x1080 closed GOP of 15 pictures, 30 DV frames (720x576
    4:2:0 and 720x480 4:1:1 alternating), 12 H.264 pictures (4 each of 1280x720, 1920x1088 and
    3840x2160, 4 resident references per size, weighted bi-prediction, deblocking) -- where every
    frame is reconstructed, converted (BT.709 limited YUV->RGBA) and composited (3 premultiplied
    RGBA8 layers, alpha-over, at its output size; conversion and composite as one pass,
    mmr_recon_yuv_composite_submit). A step is one batch per codec, then a conversion + composite
    batch: the chain of the 3840x2160 H.264 pictures on the context's stream, every other chain
    (MJPEG, MPEG-2, DV, the smaller H.264 pictures) on a second context's stream beside it (the
    deblocking wavefront of the 3840x2160 pictures is bound by its dependency chain and leaves most
    compute units idle). The plain path runs the same chain one frame and one launch at a time on
    one stream. Mpix count output frames."""
    name = "mixed"
    codec = None
    data = "synthetic (seeded post-entropy buffers per codec, seeded premultiplied RGBA8 layers)"
    metric = "reconstructed + composited Mpix/s, mixed codecs (batched vs plain one-frame-per-launch)"
    metric_unit = "Mpix/s"
    kernel = "all kernels of the step"
    streams = 2
    dtype = "u8/i16 in, int32 + f64 (IDCT) arithmetic"
    H264_SIZES = ((80, 45), (120, 68), (240, 135))

    def __init__(self, ctx, seed=2024):
        self.ctx = ctx
        self.side = mmrecode_amd.Context(device=0)   # second stream; buffers are the device's, shared
        t = time.time()
        self.frames = []   # (codec, unit, device unit)
        for i in range(30):
            u = wl.mjpeg_frame(seed, 2000 + i, rgb=False)
            self.frames.append((abi.CODEC_MJPEG, u, wl.DeviceUnit(ctx, u)))
        gop = wl.mpeg2_gop(seed, 77)
        dev = []
        for i, u in enumerate(gop):
            extra = {"forward_source": -1 if u.meta["fwd"] is None else u.meta["fwd"],
                     "backward_source": -1 if u.meta["bwd"] is None else u.meta["bwd"]}
            for tag in ("fwd", "bwd"):
                if u.meta[tag] is not None:
                    for n in ("y", "cb", "cr"):
                        extra[f"{tag}_{n}"] = dev[u.meta[tag]].ptr(n)
            dev.append(wl.DeviceUnit(ctx, u, extra))
            self.frames.append((abi.CODEC_MPEG2, u, dev[-1]))
        self.mpeg2_gop = gop
        for i in range(30):
            u = wl.dv_frame(seed, 2000 + i, system=i % 2)
            self.frames.append((abi.CODEC_DV, u, wl.DeviceUnit(ctx, u)))
        self.h264_host_refs = {}
        self.h264_dev_refs = {}
        for wmb, hmb in self.H264_SIZES:
            host = [dict(zip(("y", "cb", "cr"), wl.h264_reference_planes(seed, wmb, hmb, 50 + i))) for i in range(4)]
            devr = [{n: ctx.upload(a) for n, a in r.items()} for r in host]
            self.h264_host_refs[(wmb, hmb)] = host
            self.h264_dev_refs[(wmb, hmb)] = devr
            ptrs = [{n: b.ptr for n, b in r.items()} for r in devr]
            for i in range(4):
                u = wl.h264_picture(seed, 200 + i, wmb, hmb)
                self.frames.append((abi.CODEC_H264, u, wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, ptrs, "dev"))))
        # conversion + composite of every frame: its RGBA target and 3 layers per output geometry
        rng = np.random.Generator(np.random.PCG64([seed, 88]))
        self.layers, self.rgba, self.convert, self.composite = {}, [], [], []
        self.mpix_list = []
        self.bytes = 0
        for codec, u, d in self.frames:
            w, h, cw, ch, ls, cs, names = ServedMix.geometry(codec, u)
            if (w, h) not in self.layers:
                bufs = []
                for _ in range(3):
                    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
                    px[..., :3] = (px[..., :3].astype(np.uint32) * px[..., 3:4] // 255).astype(np.uint8)
                    bufs.append((px, ctx.upload(px)))
                self.layers[(w, h)] = bufs
            out = ctx.alloc(w * h * 4)
            self.rgba.append(out)
            conv = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=ls,
                           chroma_stride=cs, range=abi.RANGE_LIMITED, matrix=abi.MATRIX_BT709, rgba=out.ptr)
            for k, n in enumerate(names):
                conv.planes[k] = d.ptr(n)
            comp = abi.new(abi.CompositeJob, target=out.ptr, width=w, height=h, stride=w * 4, layer_count=3)
            for i, (_, b) in enumerate(self.layers[(w, h)]):
                comp.layers[i] = abi.Layer(rgba=b.ptr, width=w, height=h, stride=w * 4, x=0, y=0,
                                           op=abi.BLEND_PREMULTIPLIED_OVER, opacity=255)
            self.convert.append(conv)
            self.composite.append(comp)
            self.mpix_list.append(metrics.mpix(w, h))
            decode = {abi.CODEC_MJPEG: metrics.mjpeg_frame_bytes, abi.CODEC_MPEG2: metrics.mpeg2_picture_bytes,
                      abi.CODEC_DV: metrics.dv_frame_bytes, abi.CODEC_H264: metrics.h264_picture_bytes}[codec](u)
            # conversion + composite in one pass: planes read, 3 layers read, the target written once
            self.bytes += decode + (w * h + 2 * cw * ch) + 16 * w * h
        self.by_codec = {c: [d.struct for cc, _, d in self.frames if cc == c]
                         for c in (abi.CODEC_MJPEG, abi.CODEC_MPEG2, abi.CODEC_DV, abi.CODEC_H264)}
        self.submitters = {abi.CODEC_MJPEG: ctx.submit_mjpeg, abi.CODEC_MPEG2: ctx.submit_mpeg2, abi.CODEC_DV: ctx.submit_dv,
                           abi.CODEC_H264: ctx.submit_h264}
        # The 3840x2160 H.264 pictures (the longest deblocking chains) on the context's stream,
        # every other frame on the second stream.
        big = self.H264_SIZES[-1]

        def on_main(c, u):
            return c == abi.CODEC_H264 and (u.meta["width_mbs"], u.meta["height_mbs"]) == big
        main = [i for i, (c, u, _) in enumerate(self.frames) if on_main(c, u)]
        side = [i for i, (c, u, _) in enumerate(self.frames) if not on_main(c, u)]
        self.h264_main = [self.frames[i][2].struct for i in main]
        self.h264_side = [d.struct for c, u, d in self.frames if c == abi.CODEC_H264 and not on_main(c, u)]
        self.chains = {"side": ([self.convert[i] for i in side], [self.composite[i] for i in side]),
                       "main": ([self.convert[i] for i in main], [self.composite[i] for i in main])}
        self.pictures = len(self.frames)
        self.mpix = sum(self.mpix_list)
        self.f64_ops = None
        self.step_tickets = {}
        log(f"mixed workload: {self.pictures} frames generated and uploaded in {time.time() - t:.1f}s")

    STAGES = ("mjpeg", "mpeg2", "dv", "h264", "yuv_composite")

    def submit(self, flags=0):
        """One step; returns the ticket of the H.264 chain's last submission (wait() waits for the
        other stream's chain as well)."""
        c = self.ctx
        h = [c.submit_h264(self.h264_main, flags)]
        h.append(c.submit_yuv_composite(*self.chains["main"], flags))
        s = self.side
        o = [s.submit_mjpeg(self.by_codec[abi.CODEC_MJPEG], flags), s.submit_mpeg2(self.by_codec[abi.CODEC_MPEG2], flags),
             s.submit_dv(self.by_codec[abi.CODEC_DV], flags), s.submit_h264(self.h264_side, flags)]
        o.append(s.submit_yuv_composite(*self.chains["side"], flags))
        self.step_tickets[h[-1]] = (h, o)
        return h[-1]

    def close(self):
        self.side.close()

    def autotune(self):
        """Each context's launch shapes tuned on its own batches of the step: H.264 (the 3840x2160
        pictures) on the main context; MJPEG, MPEG-2, DV and the smaller H.264 pictures on the
        second. Returns {"main": tuning, "side": tuning}."""
        main, _ = self.ctx.autotune(abi.CODEC_H264, self.h264_main)
        side = None
        for codec, items in ((abi.CODEC_MJPEG, self.by_codec[abi.CODEC_MJPEG]), (abi.CODEC_MPEG2, self.by_codec[abi.CODEC_MPEG2]),
                             (abi.CODEC_DV, self.by_codec[abi.CODEC_DV]), (abi.CODEC_H264, self.h264_side)):
            side, _ = self.side.autotune(codec, items)
        return {"main": main, "side": side}

    def wait(self, ticket):
        self.ctx.wait(ticket)
        self.side.wait(self.step_tickets[ticket][1][-1])

    def synchronize(self):
        self.ctx.synchronize()
        self.side.synchronize()

    def step_phases(self, tickets):
        """Device ms of each stage of a step (the library's events around each submission's
        kernels, on either stream; the streams overlap), averaged over the given steps."""
        def ms(t, stage):
            h, o = self.step_tickets[t]
            pick = {"mjpeg": [(self.side, o[0])], "mpeg2": [(self.side, o[1])], "dv": [(self.side, o[2])],
                    "h264": [(self.ctx, h[0]), (self.side, o[3])], "yuv_composite": [(self.ctx, h[1]), (self.side, o[4])]}[stage]
            return sum(cx.kernel_ms(k) for cx, k in pick)
        return {name: float(np.mean([ms(t, name) for t in tickets])) for name in self.STAGES}

    def plain_items(self):
        return list(range(len(self.frames)))

    def submit_items(self, i, flags):
        codec, _, d = self.frames[i]
        self.ctx.wait(self.submitters[codec]([d.struct], flags))
        return self.ctx.submit_yuv_composite([self.convert[i]], [self.composite[i]], flags)

    def config(self):
        return {"workload": "mixed codecs per step: 30 MJPEG 640x480, one MPEG-2 1920x1080 GOP of 15, 30 DV "
                            "720x576/720x480, 12 H.264 1280x720/1920x1088/3840x2160 pictures; every frame + BT.709 "
                            "limited YUV->RGBA + 3-layer premultiplied alpha-over at its output size (one pass)",
                "pictures_per_step": self.pictures, "submissions_per_step": 7}

    def oracle_chain(self, L, i, o, mpeg2_host):
        """Frame i through the oracle: reconstruction into o (its host planes), BT.709 limited
        YUV->RGBA, the 3-layer composite; returns the RGBA frame. MPEG-2 pictures must come in
        decode order with one mpeg2_host dict (GOP index -> planes) for the GOP."""
        codec, u, _ = self.frames[i]
        if codec == abi.CODEC_MPEG2:
            mpeg2_host[self.mpeg2_gop.index(u)] = o
            assert L.oracle_mpeg2_reconstruct(C.byref(u.host_struct(o, wl.mpeg2_ref_ptrs(mpeg2_host, u, "host")))) == 0
        elif codec == abi.CODEC_H264:
            refs = self.h264_host_refs[(u.meta["width_mbs"], u.meta["height_mbs"])]
            assert L.oracle_h264_reconstruct(C.byref(u.host_struct(o, wl.h264_ref_ptrs(u, refs, "host")))) == 0
        else:
            fn = L.oracle_mjpeg_reconstruct if codec == abi.CODEC_MJPEG else L.oracle_dv_reconstruct
            assert fn(C.byref(u.host_struct(o))) == 0
        w, h, cw, ch, ls, cs, names = ServedMix.geometry(codec, u)
        rgba = np.empty((h, w, 4), np.uint8)
        conv = abi.new(abi.YuvToRgbJob, width=w, height=h, chroma_width=cw, chroma_height=ch, luma_stride=ls,
                       chroma_stride=cs, range=abi.RANGE_LIMITED, matrix=abi.MATRIX_BT709, rgba=rgba.ctypes.data)
        for k, nm in enumerate(names):
            conv.planes[k] = o[nm].ctypes.data
        assert L.oracle_yuv_to_rgb(C.byref(conv)) == 0
        comp = abi.new(abi.CompositeJob, target=rgba.ctypes.data, width=w, height=h, stride=w * 4, layer_count=3)
        for k, (px, _) in enumerate(self.layers[(w, h)]):
            comp.layers[k] = abi.Layer(rgba=px.ctypes.data, width=w, height=h, stride=w * 4, x=0, y=0,
                                       op=abi.BLEND_PREMULTIPLIED_OVER, opacity=255)
        assert L.oracle_composite(C.byref(comp)) == 0
        return rgba

    def cpu_sample(self, threads):
        """The whole step's frames through the oracle chain (reconstruction, conversion, composite),
        spread over the threads (the MPEG-2 GOP on one thread, in decode order)."""
        from oracle import binding

        L = binding.lib()
        n = max(1, threads)
        items = []
        for i, (codec, u, _) in enumerate(self.frames):
            if codec != abi.CODEC_MPEG2:
                items.append([i])
        items.insert(0, [i for i, (c, _, _) in enumerate(self.frames) if c == abi.CODEC_MPEG2])
        groups = [[] for _ in range(n)]
        for k, it in enumerate(items):
            groups[k % n] += it
        outs = {i: u.host_outputs() for i, (_, u, _) in enumerate(self.frames)}
        mpeg2_host = {}
        chain = lambda i: self.oracle_chain(L, i, outs[i], mpeg2_host)
        jobs = [lambda g=g: [chain(i) for i in g] for g in groups if g]
        passes, dt = cpu_passes(jobs)
        return (passes * self.mpix / dt, len(jobs),
                f"{passes} pass(es) over the step's {self.pictures} frames (reconstruction + YUV->RGBA + composite), "
                f"spread over {len(jobs)} threads (the MPEG-2 GOP on one), {dt:.1f} s wall")


class LinearWorkload:
    """MMFX linear-light compositing (LinearPixel::over, crates/mmfx/src/render.rs:1826-1854): 16
    targets of 1920x1080 premultiplied linear u16 RGBA per launch, each with a full-frame surface at
    group opacity 0.5 (Surface::blend_surface), a 1280x720 sRGBA8 image at (320, 180)
    (execute_draw_image) and a second full-frame surface at a random opacity."""
    name = "linear"
    codec = None
    data = "synthetic (seeded premultiplied linear surfaces and sRGBA8 images)"
    metric = "composited linear-light Mpix/s (u16 premultiplied RGBA, 3 layers per target)"
    metric_unit = "Mpix/s"
    kernel = "linear_composite_kernel"
    dtype = "u16 / u8 in, u16 out, int32 arithmetic"

    def __init__(self, ctx, frames=16, width=1920, height=1080, seed=2024):
        self.ctx, self.width, self.height = ctx, width, height
        rng = np.random.default_rng(seed)
        self.host, self.bufs, self.structs = [], [], []
        for _ in range(frames):
            target = self.surface(rng, height, width)
            s1, s2 = self.surface(rng, height, width), self.surface(rng, height, width)
            img = rng.integers(0, 256, (720, 1280, 4), dtype=np.uint8)
            layers = [(s1, abi.LINEAR_SOURCE_SURFACE, 0, 0, 32768), (img, abi.LINEAR_SOURCE_SRGBA8, 320, 180, 65535),
                      (s2, abi.LINEAR_SOURCE_SURFACE, 0, 0, int(rng.integers(1, 65535)))]
            dt = ctx.upload(target)
            dl = [ctx.upload(a) for a, *_ in layers]
            self.structs.append(self.job(dt.ptr, layers, [b.ptr for b in dl]))
            self.host.append((target, layers))
            self.bufs += [dt] + dl
        self.pictures = frames
        self.mpix = frames * metrics.mpix(width, height)
        # target read + written (16 B), two full-frame surfaces (16 B) per pixel, the image (4 B per image pixel)
        self.bytes = frames * (width * height * 32 + 1280 * 720 * 4)
        self.f64_ops = None

    @staticmethod
    def surface(rng, h, w):
        a = rng.integers(0, 65536, (h, w, 1), dtype=np.uint32)
        c = rng.integers(0, 65536, (h, w, 3), dtype=np.uint32) * a // 65535
        return np.concatenate([c, a], axis=2).astype(np.uint16)

    def job(self, target, layers, ptrs):
        j = abi.new(abi.LinearCompositeJob, target=target, width=self.width, height=self.height, stride=self.width,
                    layer_count=len(layers))
        for i, ((a, fmt, x, y, op), p) in enumerate(zip(layers, ptrs)):
            j.layers[i] = abi.LinearLayer(pixels=p, format=fmt, stride=a.shape[1], x=x, y=y, width=a.shape[1], height=a.shape[0],
                                          clip_left=0, clip_top=0, clip_right=self.width, clip_bottom=self.height, opacity=op)
        return j

    def submit(self, flags=0):
        return self.ctx.submit_linear_composite(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_linear_composite(items, flags)

    def config(self):
        return {"workload": "MMFX linear-light composite, 16 targets 1920x1080 x 3 layers per launch",
                "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        n = max(1, min(threads, len(self.host)))
        jobs = []
        for target, layers in self.host[:n]:
            t = target.copy()
            jobs.append((self.job(t.ctypes.data, layers, [a.ctypes.data for a, *_ in layers]), t))
        passes, dt = cpu_passes([lambda j=j: L.oracle_linear_composite(C.byref(j)) for j, _ in jobs])
        return (passes * n * metrics.mpix(self.width, self.height) / dt, n,
                f"{passes} pass(es) over {n} targets 1920x1080 x 3 layers, one target per thread, {dt:.1f} s wall on {n} threads")


WORKLOADS = {"mpeg2": Mpeg2Workload, "mjpeg": MjpegWorkload, "dv": DvWorkload, "h264": H264Workload, "mixed": MixedWorkload,
             "h264i": H264IntraWorkload, "scale": ScaleWorkload, "overlay": OverlayWorkload, "linear": LinearWorkload}


def calibrate_copy(nbytes=1 << 30, repeats=5):
    """Device-to-device copies of a known size (the HIP runtime's copy kernel), for scaling the HBM
    byte counters of rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE."""
    ctx = mmrecode_amd.Context(device=0)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = C.c_int
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.restype = C.c_int
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.write(src, np.ones(1 << 20, np.uint8))
    hip_memcpy_device_to_device = 3
    for _ in range(repeats):
        assert hip.hipMemcpy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, hip_memcpy_device_to_device) == 0
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"calibration": "device-to-device copy", "bytes_per_copy": nbytes, "copies": repeats}), flush=True)
    ctx.close()


def phase_names(w, n):
    """Names of the library's launch phases of one batched submission of workload w."""
    if hasattr(w, "phase_names"):
        return w.phase_names(n)
    names = getattr(w, "phases", (w.kernel,))
    return tuple(names[:n]) if n <= len(names) else tuple(names) + tuple(f"phase{i}" for i in range(len(names), n))


def measured_traffic(workload):
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f).get(workload)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mpeg2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain one-picture-per-launch leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes instead of autotuning them")
    ap.add_argument("--served", action="store_true", help="also run the served mixed-workload sweep")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run device-to-device copies of 1 GiB (HBM byte-counter calibration)")
    ap.add_argument("--served-rates", default="50,200,500,1000,2000")
    ap.add_argument("--served-seconds", type=float, default=2.0)
    ap.add_argument("--plain-seconds", type=float, default=4.0,
                    help="wall-time cap of each plain replay of a served trace (0: skip)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    ctx = mmrecode_amd.Context(device=0)
    w = WORKLOADS[args.workload](ctx)
    log(f"{w.pictures} pictures, {w.mpix:.1f} Mpix, {w.bytes / 1e9:.3f} GB algorithmic per step")

    tuning = {"chosen": ctx.get_tuning(), "trials": None}
    if not args.no_autotune and w.codec is not None:
        t0 = time.perf_counter()
        best, trials = ctx.autotune(w.codec, w.structs, repeats=3)
        tuning = {"chosen": best, "trials": [{"tuning": t, "kernel_ms": round(ms, 4)} for t, ms in trials],
                  "seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotuned in {tuning['seconds']} s: {best}")
    elif not args.no_autotune and hasattr(w, "autotune"):   # several codecs / contexts: the workload tunes each
        t0 = time.perf_counter()
        tuning = {"chosen": w.autotune(), "trials": None}
        tuning["seconds"] = round(time.perf_counter() - t0, 2)
        log(f"autotuned in {tuning['seconds']} s: {tuning['chosen']}")
    wait = getattr(w, "wait", ctx.wait)                  # a workload on two streams waits for both
    sync = getattr(w, "synchronize", ctx.synchronize)
    for _ in range(args.warmup):
        wait(w.submit())
    sync()
    # Frame latency of the batched path: one batch in flight, submit -> completion observed on the
    # host; every picture of a batch completes with it.
    batch_lat, idle_submit = [], []
    for _ in range(args.steps):
        ts = time.perf_counter()
        t = w.submit()
        idle_submit.append(time.perf_counter() - ts)
        wait(t)
        batch_lat.append((time.perf_counter() - ts) * 1e3)
    t0 = time.perf_counter()
    tickets, submit_s = [], []
    for _ in range(args.steps):
        ts = time.perf_counter()
        tickets.append(w.submit())
        submit_s.append(time.perf_counter() - ts)
    for t in tickets:
        wait(t)
    elapsed = time.perf_counter() - t0
    # Device time of each launch phase (HIP events the library records on the submission stream
    # between phases; they add up to the submission's kernel span).
    if hasattr(w, "step_phases"):   # several submissions per step: the device time of each
        phase_ms = w.step_phases(tickets)
        span_ms = float(sum(phase_ms.values()))
    else:
        span_ms = float(np.mean([ctx.kernel_ms(t) for t in tickets]))
        phase_lists = [ctx.kernel_phases_ms(t) for t in tickets]
        names = phase_names(w, len(phase_lists[0]))
        phase_ms = {n: float(np.mean([p[i] for p in phase_lists])) for i, n in enumerate(names)}
    k_ms = phase_ms.get(w.kernel, span_ms)
    ms_per_step = elapsed * 1e3 / args.steps
    if getattr(w, "streams", 1) > 1 and w.kernel not in phase_ms:
        k_ms = ms_per_step   # overlapping streams: the phases do not add up; the step's wall time bounds the span
    value = w.mpix * args.steps / elapsed
    log(f"batched: {ms_per_step:.3f} ms/step, {value:.1f} Mpix/s, kernels {span_ms:.4f} ms "
        f"({', '.join(f'{n} {v:.4f}' for n, v in phase_ms.items())})")

    # The same batch prepared once (job records resident, launches replayed from a HIP graph).
    prepared = None
    if w.codec is not None:
        prep = ctx.prepare(w.codec, w.structs)
        for _ in range(max(1, args.warmup)):
            ctx.wait(prep.launch())
        ctx.synchronize()
        t0 = time.perf_counter()
        tickets = [prep.launch() for _ in range(args.steps)]
        for t in tickets:
            ctx.wait(t)
        p_elapsed = time.perf_counter() - t0
        p_kernel_ms = float(np.mean([ctx.kernel_ms(t) for t in tickets]))
        prep.close()
        prepared = {"value": round(w.mpix * args.steps / p_elapsed, 3), "unit": "Mpix/s",
                    "ms_per_step": round(p_elapsed * 1e3 / args.steps, 4), "kernel_ms": round(p_kernel_ms, 4),
                    "mode": "batch prepared once, launches replayed from a HIP graph"}
        log(f"prepared graph: {prepared['ms_per_step']:.3f} ms/step, {prepared['value']:.1f} Mpix/s")

    plain = None
    latencies = []
    if not args.no_plain:
        items = w.plain_items()
        ctx.wait(w.submit_items(items[0], abi.SUBMIT_PER_FRAME_LAUNCH))
        t0 = time.perf_counter()
        for it in items:
            ts = time.perf_counter()
            ctx.wait(w.submit_items(it, abi.SUBMIT_PER_FRAME_LAUNCH))
            latencies.append((time.perf_counter() - ts) * 1e3)
        dt = time.perf_counter() - t0
        plain = w.mpix / dt
        log(f"plain one-picture-per-launch: {plain:.1f} Mpix/s, p50 {percentile(latencies, 50):.3f} ms")

    cpu = None
    if not args.no_cpu:
        usable, nproc, model = host_cpu()
        # every thread this process may use, up to the host share of one GPU on these boxes (16)
        threads = args.cpu_threads or min(16, usable)
        v, cores, sample = w.cpu_sample(threads)
        cpu = {"value": round(v, 3), "unit": "Mpix/s", "cores": cores, "kind": "port", "sample": sample,
               "host": {"nproc": nproc, "usable_threads": usable, "cpu_model": model}}
        log(f"cpu baseline: {v:.2f} Mpix/s on {cores} threads")

    served = None
    if args.served:
        rates = [float(r) for r in args.served_rates.split(",") if r]
        served = run_served_sweep(ctx, rates, args.served_seconds, args.plain_seconds)

    kernel_bytes = getattr(w, "kernel_bytes", {w.kernel: w.bytes})
    achieved = kernel_bytes[w.kernel] / (k_ms * 1e-3) / 1e9 if k_ms else None
    phases = []
    for n, ms in phase_ms.items():
        ph = {"kernel": n, "ms": round(ms, 4)}
        if n in kernel_bytes and ms > 0:
            ph["algorithmic_bytes"] = int(kernel_bytes[n])
            ph["achieved_gbs"] = round(kernel_bytes[n] / (ms * 1e-3) / 1e9, 2)
        phases.append(ph)
    result = {
        "metric": getattr(w, "metric", "reconstructed Mpix/s batched-served vs plain GPU vs CPU; p50/p99 frame latency"),
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": getattr(w, "dtype", "u8/i16 in, int32 + f64 (IDCT) arithmetic"),
        "data": getattr(w, "data", "synthetic (seeded post-entropy coefficient / macroblock buffers)"),
        "config": w.config(),
        "plain_gpu": {"value": round(plain, 3) if plain else None, "unit": "Mpix/s",
                      "p50_frame_ms": percentile(latencies, 50), "p99_frame_ms": percentile(latencies, 99),
                      "mode": "one picture per launch, synchronous"},
        "batched_frame_latency_ms": {"p50": round(percentile(batch_lat, 50), 4), "p99": round(percentile(batch_lat, 99), 4),
                                     "samples": len(batch_lat),
                                     "note": "submit -> completion of one batch with nothing else in flight; every "
                                             "picture of a batch completes with it"},
        "speedup_vs_plain": round(value / plain, 2) if plain else None,
        "host_submit_ms": {"pipelined": round(float(np.median(submit_s)) * 1e3, 4),
                           "device_idle": round(float(np.median(idle_submit)) * 1e3, 4),
                           "note": "median wall time of one submit call (host work + any wait inside it)"},
        "prepared_graph": prepared,
        "tuning": tuning,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": None, "kernel": w.kernel, "kernel_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": int(kernel_bytes[w.kernel]),
                     "per": ("step: every launch of the step on both streams (bytes summed; time = the step's wall time, "
                             "the streams overlap)" if getattr(w, "streams", 1) > 1 and w.kernel not in phase_ms else
                             "step: every launch of the kernel in one step (bytes and device time summed over them)"),
                     "phases": phases, "kernels_ms_per_step": round(span_ms, 4),
                     "timing": "HIP events recorded by the library on the submission stream around each launch phase"},
        "cpu_baseline": cpu,
        "served": served,
    }
    tr = measured_traffic(args.workload)
    if tr and tr.get("kernel") == w.kernel:
        result["roofline"]["traffic"] = int(tr["hbm_bytes_per_step"])
        result["roofline"]["traffic_source"] = tr["source"]
        result["roofline"]["launches_per_step"] = tr["dispatches_per_step"]
    if w.f64_ops and k_ms:
        result["roofline"]["fp64_dense_tflops_equiv"] = round(w.f64_ops / (k_ms * 1e-3) / 1e12, 3)
        result["roofline"]["fp64_peak_tflops"] = FP64_PEAK_TFLOPS
    print(json.dumps(result), flush=True)
    if hasattr(w, "close"):
        w.close()
    ctx.close()


if __name__ == "__main__":
    main()
