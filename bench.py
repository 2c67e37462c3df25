"""Benchmark of the MI355X reconstruction path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mpeg2|mjpeg|dv|h264]

A step is one pass of the hot path over one batch of synthetic post-entropy input, all inputs
resident in HBM before the timed region:
  mpeg2 (default)  8 closed GOPs x 15 pictures of 1920x1080 4:2:0 in one persistent launch
  mjpeg            120 frames of 640x480 4:2:0 + RGB8 in one launch
  dv               256 frames, 720x576 4:2:0 and 720x480 4:1:1 interleaved, one launch
  h264             16 pictures each of 1280x720, 1920x1080, 3840x2160 with deblocking, one launch
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
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(values, q):
    return float(np.percentile(np.asarray(values, float), q)) if len(values) else None


# ------------------------------------------------------------------ workloads
class Mpeg2Workload:
    name = "mpeg2"
    metric_unit = "Mpix/s"
    kernel = "mpeg2_kernel"

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
                            f"pictures per launch (configs[1])", "pictures_per_step": self.pictures,
                "resolution": f"{self.width}x{self.height}", "gop": wl.GOP_DECODE_ORDER}

    def cpu_sample(self, threads):
        """One GOP per thread through the oracle (threads <= GOPs), Mpix/s."""
        from oracle import binding

        L = binding.lib()
        n = max(1, min(threads, len(self.gops)))

        def run_gop(g, out_mpix):
            outs = {}
            for i, u in enumerate(self.gops[g]):
                outs[i] = u.host_outputs()
                s = u.host_struct(outs[i], wl.mpeg2_ref_ptrs(outs, u, "host"))
                assert L.oracle_mpeg2_reconstruct(C.byref(s)) == 0
            out_mpix.append(len(self.gops[g]) * metrics.mpix(self.width, self.height))

        done = []
        ts = [threading.Thread(target=run_gop, args=(g, done)) for g in range(n)]
        t = time.time()
        for th in ts:
            th.start()
        for th in ts:
            th.join()
        dt = time.time() - t
        return sum(done) / dt, n, f"{n} GOP(s) x 15 pictures {self.width}x{self.height}, one GOP per thread, {dt:.1f}s"


class MjpegWorkload:
    name = "mjpeg"
    metric_unit = "Mpix/s"
    kernel = "mjpeg_blocks_kernel"

    def __init__(self, ctx, frames=120, seed=2024):
        self.ctx = ctx
        self.units = [wl.mjpeg_frame(seed, i) for i in range(frames)]
        self.dev = [wl.DeviceUnit(ctx, u) for u in self.units]
        self.structs = [d.struct for d in self.dev]
        self.pictures = frames
        self.mpix = frames * metrics.mpix(640, 480)
        self.bytes = sum(metrics.mjpeg_frame_bytes(u) for u in self.units)
        self.f64_ops = sum(u.meta["blocks"] for u in self.units) * 2048

    def submit(self, flags=0):
        return self.ctx.submit_mjpeg(self.structs, flags)

    def plain_items(self):
        return [[s] for s in self.structs]

    def submit_items(self, items, flags):
        return self.ctx.submit_mjpeg(items, flags)

    def config(self):
        return {"workload": "MJPEG baseline 640x480 4:2:0 + BT.601 RGB8, 120 frames per launch (configs[0] frames)",
                "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        n = max(1, min(threads, len(self.units)))
        outs = [u.host_outputs() for u in self.units[:n]]
        t = time.time()
        ts = [threading.Thread(target=lambda u=u, o=o: L.oracle_mjpeg_reconstruct(C.byref(u.host_struct(o))))
              for u, o in zip(self.units[:n], outs)]
        for th in ts:
            th.start()
        for th in ts:
            th.join()
        dt = time.time() - t
        return n * metrics.mpix(640, 480) / dt, n, f"{n} frames 640x480, one frame per thread, {dt:.2f}s"


class DvWorkload:
    name = "dv"
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
        return {"workload": "DV25 720x576 4:2:0 + 720x480 4:1:1 interleaved, 256 frames per launch (configs[2])",
                "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        n = max(1, min(threads, len(self.units)))
        outs = [u.host_outputs() for u in self.units[:n]]
        t = time.time()
        ts = [threading.Thread(target=lambda u=u, o=o: L.oracle_dv_reconstruct(C.byref(u.host_struct(o))))
              for u, o in zip(self.units[:n], outs)]
        for th in ts:
            th.start()
        for th in ts:
            th.join()
        dt = time.time() - t
        mp = sum(metrics.mpix(720, u.meta["height"]) for u in self.units[:n])
        return mp / dt, n, f"{n} DV frames, one frame per thread, {dt:.2f}s"


class H264Workload:
    name = "h264"
    metric_unit = "Mpix/s"
    kernel = "h264_mb_kernel"
    SIZES = ((80, 45), (120, 68), (240, 135))

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
                u = wl.h264_picture(seed, i, wmb, hmb)
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
                            "3840x2160 per launch (configs[3])", "pictures_per_step": self.pictures}

    def cpu_sample(self, threads):
        from oracle import binding

        L = binding.lib()
        picks = [u for u in self.units if u.meta["width_mbs"] == 120][: max(1, min(threads, 16))]
        outs = [u.host_outputs() for u in picks]
        t = time.time()

        def run(u, o):
            s = u.host_struct(o, wl.h264_ref_ptrs(u, self.host_refs[(u.meta["width_mbs"], u.meta["height_mbs"])], "host"))
            assert L.oracle_h264_reconstruct(C.byref(s)) == 0

        ts = [threading.Thread(target=run, args=(u, o)) for u, o in zip(picks, outs)]
        for th in ts:
            th.start()
        for th in ts:
            th.join()
        dt = time.time() - t
        mp = len(picks) * metrics.mpix(1920, 1088)
        return mp / dt, len(picks), f"{len(picks)} 1920x1088 pictures, one picture per thread, {dt:.2f}s"


WORKLOADS = {"mpeg2": Mpeg2Workload, "mjpeg": MjpegWorkload, "dv": DvWorkload, "h264": H264Workload}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mpeg2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain one-picture-per-launch leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    ctx = mmrecode_amd.Context(device=0)
    w = WORKLOADS[args.workload](ctx)
    log(f"{w.pictures} pictures, {w.mpix:.1f} Mpix, {w.bytes / 1e9:.3f} GB algorithmic per step")

    for _ in range(args.warmup):
        ctx.wait(w.submit())
    ctx.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    tickets = [w.submit() for _ in range(args.steps)]
    for t in tickets:
        ctx.wait(t)
    elapsed = time.perf_counter() - t0
    for t in tickets:
        kernel_ms.append(ctx.kernel_ms(t))
    ms_per_step = elapsed * 1e3 / args.steps
    value = w.mpix * args.steps / elapsed
    k_ms = float(np.mean(kernel_ms)) if kernel_ms else None
    log(f"batched: {ms_per_step:.3f} ms/step, {value:.1f} Mpix/s, kernel {k_ms} ms")

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
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, cores, sample = w.cpu_sample(threads)
        cpu = {"value": round(v, 3), "unit": "Mpix/s", "cores": cores, "kind": "port", "sample": sample}
        log(f"cpu baseline: {v:.2f} Mpix/s on {cores} threads")

    achieved = w.bytes / (k_ms * 1e-3) / 1e9 if k_ms else None
    result = {
        "metric": "reconstructed Mpix/s batched-served vs plain GPU vs CPU; p50/p99 frame latency",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/i16 in, int32 + f64 (IDCT) arithmetic",
        "data": "synthetic (seeded post-entropy coefficient / macroblock buffers)",
        "config": w.config(),
        "plain_gpu": {"value": round(plain, 3) if plain else None, "unit": "Mpix/s",
                      "p50_frame_ms": percentile(latencies, 50), "p99_frame_ms": percentile(latencies, 99),
                      "mode": "one picture per launch, synchronous"},
        "batched_frame_latency_ms": {"p50": round(ms_per_step, 4), "p99": round(ms_per_step, 4),
                                     "note": "every picture of a batch completes with its launch"},
        "speedup_vs_plain": round(value / plain, 2) if plain else None,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": None, "kernel": w.kernel, "kernel_ms": k_ms,
                     "algorithmic_bytes_per_launch": int(w.bytes)},
        "cpu_baseline": cpu,
    }
    if w.f64_ops and k_ms:
        result["roofline"]["fp64_dense_tflops_equiv"] = round(w.f64_ops / (k_ms * 1e-3) / 1e12, 3)
        result["roofline"]["fp64_peak_tflops"] = FP64_PEAK_TFLOPS
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
