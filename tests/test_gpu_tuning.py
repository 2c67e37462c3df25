"""Launch-shape tuning: every setting reconstructs the same bytes as the oracle, and the autotuner
keeps the fastest candidate it timed."""
import numpy as np
import pytest

import mmrecode_amd
from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

from test_gpu_parity import _assert_same, _h264_refs, _mpeg2_gpu, _mpeg2_oracle, _oracle_run

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuned(ctx):
    """Restores the context's tuning after the test."""
    saved = ctx.get_tuning()
    yield ctx
    ctx.set_tuning(**saved)


@pytest.mark.parametrize("blocks", [16, 64])
def test_block_kernels_every_group_size_bit_exact(tuned, oracle, blocks):
    ctx = tuned
    ctx.set_tuning(mjpeg_group_blocks=blocks, mpeg2_group_blocks=blocks, dv_group_blocks=blocks)
    # MJPEG: ragged 4:2:0 plus a full frame
    for u in (wl.mjpeg_frame(seed=7, index=blocks, width=45, height=29), wl.mjpeg_frame(seed=7, index=1)):
        d = wl.DeviceUnit(ctx, u)
        ctx.wait(ctx.submit_mjpeg([d.struct]))
        want = _oracle_run(oracle, "oracle_mjpeg_reconstruct", u)
        for n in want:
            _assert_same(f"mjpeg {blocks} {n}", d.download(n), want[n])
        d.free()
    # MPEG-2 GOP (dependency levels), both launch modes
    units = wl.mpeg2_gop(seed=29, gop_index=blocks, width=176, height=144)
    want = _mpeg2_oracle(oracle, units)
    for flags in (0, abi.SUBMIT_PER_FRAME_LAUNCH):
        dev = _mpeg2_gpu(ctx, units, flags)
        for i, d in enumerate(dev):
            for n in ("y", "cb", "cr"):
                _assert_same(f"mpeg2 {blocks} picture {i} {n}", d.download(n), want[i][n])
        for d in dev:
            d.free()
    # DV, both systems
    for i, system in enumerate((abi.DV_SYSTEM_525_60, abi.DV_SYSTEM_625_50)):
        u = wl.dv_frame(seed=37, index=blocks + i, system=system)
        d = wl.DeviceUnit(ctx, u)
        ctx.wait(ctx.submit_dv([d.struct]))
        want = _oracle_run(oracle, "oracle_dv_reconstruct", u)
        for n in ("y", "cb", "cr"):
            _assert_same(f"dv {blocks} {n}", d.download(n), want[n])
        d.free()


@pytest.mark.parametrize("shape", [
    dict(h264_deblock_band_rows=4, h264_deblock_groups_per_cu=1, h264_deblock_publish=1),
    dict(h264_deblock_band_rows=4, h264_deblock_groups_per_cu=2, h264_deblock_publish=8),
    dict(h264_deblock_band_rows=8, h264_deblock_groups_per_cu=4, h264_deblock_publish=2),
    dict(h264_deblock_band_rows=16, h264_deblock_groups_per_cu=2, h264_deblock_publish=32),
])
def test_h264_deblock_band_shapes_bit_exact(tuned, oracle, shape):
    ctx = tuned
    ctx.set_tuning(**shape)
    wmb, hmb = 11, 37    # 10 bands of 4 rows (the last of 1), 5 of 8, 3 of 16
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    units = [wl.h264_picture(seed=47, index=i, width_mbs=wmb, height_mbs=hmb, weighted=i == 0, deblock=True)
             for i in range(2)]
    dev = [wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, [{n: b.ptr for n, b in r.items()} for r in dev_refs], "dev"))
           for u in units]
    for flags in (0, abi.SUBMIT_PER_FRAME_LAUNCH):
        ctx.wait(ctx.submit_h264([d.struct for d in dev], flags))
        for i, (u, d) in enumerate(zip(units, dev)):
            want = _oracle_run(oracle, "oracle_h264_reconstruct", u, wl.h264_ref_ptrs(u, host_refs, "host"))
            for n in ("y", "cb", "cr"):
                _assert_same(f"{shape} flags {flags} picture {i} {n}", d.download(n), want[n])
    for d in dev:
        d.free()
    for r in dev_refs:
        for b in r.values():
            b.free()


def test_autotune_keeps_fastest_and_reconstructs(tuned, oracle):
    ctx = tuned
    units = wl.mpeg2_gop(seed=31, gop_index=1, width=352, height=288)
    dev = _mpeg2_gpu(ctx, units)
    for d in dev:   # clear the outputs the submit above wrote
        for n in ("y", "cb", "cr"):
            shape, dtype, _ = d.unit.outputs[n]
            ctx.write(d.bufs[n], np.zeros(shape, dtype))
    best, trials = ctx.autotune(abi.CODEC_MPEG2, [d.struct for d in dev], repeats=2)
    assert sorted(t["mpeg2_group_blocks"] for t, _ in trials) == [16, 32, 64]
    fastest = min(trials, key=lambda tr: tr[1])[0]
    assert best == fastest and ctx.get_tuning() == best
    want = _mpeg2_oracle(oracle, units)
    for i, d in enumerate(dev):
        for n in ("y", "cb", "cr"):
            _assert_same(f"picture {i} {n}", d.download(n), want[i][n])
    for d in dev:
        d.free()


def test_autotune_h264_candidates(tuned, oracle):
    ctx = tuned
    wmb, hmb = 8, 20
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    u = wl.h264_picture(seed=53, index=0, width_mbs=wmb, height_mbs=hmb, weighted=False, deblock=True)
    d = wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, [{n: b.ptr for n, b in r.items()} for r in dev_refs], "dev"))
    best, trials = ctx.autotune(abi.CODEC_H264, [d.struct], repeats=1)
    # band rows {4, 8, 16} x groups per compute unit {1, 2} x publish interval {1, 2, 4, 8}
    assert len(trials) == 24 and all(ms > 0 for _, ms in trials)
    assert {(t["h264_deblock_band_rows"], t["h264_deblock_groups_per_cu"], t["h264_deblock_publish"]) for t, _ in trials} == \
        {(r, g, p) for r in (4, 8, 16) for g in (1, 2) for p in (1, 2, 4, 8)}
    assert ctx.get_tuning() == best
    want = _oracle_run(oracle, "oracle_h264_reconstruct", u, wl.h264_ref_ptrs(u, host_refs, "host"))
    for n in ("y", "cb", "cr"):
        _assert_same(n, d.download(n), want[n])
    d.free()
    for r in dev_refs:
        for b in r.values():
            b.free()


def test_set_tuning_rejects_out_of_range(tuned):
    ctx = tuned
    before = ctx.get_tuning()
    for bad in (dict(mpeg2_group_blocks=48), dict(h264_deblock_band_rows=32), dict(h264_deblock_groups_per_cu=0),
                dict(h264_deblock_publish=33)):
        with pytest.raises(mmrecode_amd.ReconError) as err:
            ctx.set_tuning(**bad)
        assert err.value.status == abi.STATUS_INVALID_ARGUMENT
    assert ctx.get_tuning() == before
