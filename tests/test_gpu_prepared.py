"""Prepared batches (job records uploaded once, launches replayed from a HIP graph): every replay,
with eager submissions and synchronizations between replays and the outputs overwritten before
each one, is bit-exact against the oracle."""
import numpy as np
import pytest

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

from test_gpu_parity import _assert_same, _h264_refs, _mpeg2_oracle, _oracle_run

pytestmark = pytest.mark.gpu

REPLAYS = 3


def _poison(ctx, dev, names=("y", "cb", "cr")):
    for d in dev:
        for n in names:
            shape, dtype, _ = d.unit.outputs[n]
            ctx.write(d.bufs[n], np.full(shape, 0x5A, dtype))


def _eager_interlude(ctx, oracle, seed):
    """An ordinary MJPEG submission between replays, itself checked."""
    u = wl.mjpeg_frame(seed=seed, index=0, width=48, height=32)
    d = wl.DeviceUnit(ctx, u)
    ctx.wait(ctx.submit_mjpeg([d.struct]))
    want = _oracle_run(oracle, "oracle_mjpeg_reconstruct", u)
    for n in want:
        _assert_same(f"interlude {n}", d.download(n), want[n])
    d.free()


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_prepared_mpeg2_gop_replays(ctx, oracle, flags):
    units = wl.mpeg2_gop(seed=23, gop_index=5, width=176, height=144)
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
    want = _mpeg2_oracle(oracle, units)
    prep = ctx.prepare(abi.CODEC_MPEG2, [d.struct for d in dev], flags)
    try:
        for r in range(REPLAYS):
            _poison(ctx, dev)
            ctx.wait(prep.launch())
            for i, d in enumerate(dev):
                for n in ("y", "cb", "cr"):
                    _assert_same(f"replay {r} picture {i} {n}", d.download(n), want[i][n])
            _eager_interlude(ctx, oracle, seed=100 + r)
            ctx.synchronize()
    finally:
        prep.close()
        for d in dev:
            d.free()


@pytest.mark.parametrize("case", [
    dict(wmb=7, hmb=6, deblock=False),
    # three deblocking bands: the band tickets and hand-over counters restart at zero each replay
    dict(wmb=12, hmb=40, deblock=True),
])
def test_prepared_h264_replays(ctx, oracle, case):
    wmb, hmb = case["wmb"], case["hmb"]
    host_refs, dev_refs = _h264_refs(ctx, wmb, hmb)
    units = [wl.h264_picture(seed=43, index=i, width_mbs=wmb, height_mbs=hmb, weighted=True,
                             deblock=case["deblock"]) for i in range(2)]
    dev = [wl.DeviceUnit(ctx, u, wl.h264_ref_ptrs(u, [{n: b.ptr for n, b in r.items()} for r in dev_refs], "dev"))
           for u in units]
    want = [_oracle_run(oracle, "oracle_h264_reconstruct", u, wl.h264_ref_ptrs(u, host_refs, "host")) for u in units]
    prep = ctx.prepare(abi.CODEC_H264, [d.struct for d in dev])
    try:
        for r in range(REPLAYS):
            _poison(ctx, dev)
            ctx.wait(prep.launch())
            for i, d in enumerate(dev):
                for n in ("y", "cb", "cr"):
                    _assert_same(f"replay {r} picture {i} {n}", d.download(n), want[i][n])
            _eager_interlude(ctx, oracle, seed=200 + r)
            ctx.synchronize()
    finally:
        prep.close()
        for d in dev:
            d.free()
        for rr in dev_refs:
            for b in rr.values():
                b.free()


def test_prepared_dv_and_mjpeg_replays(ctx, oracle):
    frames = [wl.dv_frame(seed=33, index=i, system=s) for i, s in
              enumerate([abi.DV_SYSTEM_525_60, abi.DV_SYSTEM_625_50])]
    ddv = [wl.DeviceUnit(ctx, u) for u in frames]
    jpegs = [wl.mjpeg_frame(seed=13, index=i, width=80, height=48) for i in range(2)]
    djp = [wl.DeviceUnit(ctx, u) for u in jpegs]
    want_dv = [_oracle_run(oracle, "oracle_dv_reconstruct", u) for u in frames]
    want_jp = [_oracle_run(oracle, "oracle_mjpeg_reconstruct", u) for u in jpegs]
    p_dv = ctx.prepare(abi.CODEC_DV, [d.struct for d in ddv])
    p_jp = ctx.prepare(abi.CODEC_MJPEG, [d.struct for d in djp])
    try:
        for r in range(REPLAYS):
            _poison(ctx, ddv)
            _poison(ctx, djp, names=[n for n in jpegs[0].outputs])
            t1 = p_dv.launch()
            t2 = p_jp.launch()
            ctx.wait(t2)
            ctx.wait(t1)
            for i, d in enumerate(ddv):
                for n in ("y", "cb", "cr"):
                    _assert_same(f"replay {r} dv {i} {n}", d.download(n), want_dv[i][n])
            for i, d in enumerate(djp):
                for n in want_jp[i]:
                    _assert_same(f"replay {r} mjpeg {i} {n}", d.download(n), want_jp[i][n])
            ctx.synchronize()
    finally:
        p_dv.close()
        p_jp.close()
        for d in ddv + djp:
            d.free()


def test_prepare_rejects_bad_arguments(ctx):
    import mmrecode_amd

    with pytest.raises(mmrecode_amd.ReconError) as err:
        ctx.prepare(abi.CODEC_MJPEG, [])
    assert err.value.status == abi.STATUS_INVALID_ARGUMENT
