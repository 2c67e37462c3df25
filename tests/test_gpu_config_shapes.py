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
