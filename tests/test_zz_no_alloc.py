"""The allocation-free MPEG-2 entry point (mmr_recon_mpeg2_reconstruct_noalloc) against the existing
one (mmr_recon_mpeg2_submit): two calls in a row, the second with a smaller input, every plane bit
for bit equal to the existing entry point's (and to the oracle's); a submission beyond the fixed
scratch is rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

import mmrecode_amd
from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _units_on_device(ctx, units):
    dev = []
    for u in units:
        extra = {"forward_source": -1 if u.meta["fwd"] is None else u.meta["fwd"],
                 "backward_source": -1 if u.meta["bwd"] is None else u.meta["bwd"]}
        for tag in ("fwd", "bwd"):
            if u.meta[tag] is not None:
                for n in ("y", "cb", "cr"):
                    extra[f"{tag}_{n}"] = dev[u.meta[tag]].ptr(n)
        dev.append(wl.DeviceUnit(ctx, u, extra))
    return dev


def _oracle(oracle, units):
    outs = {}
    for i, u in enumerate(units):
        out = u.host_outputs()
        s = u.host_struct(out, wl.mpeg2_ref_ptrs(outs, u, "host"))
        assert oracle.oracle_mpeg2_reconstruct(C.byref(s)) == 0
        outs[i] = out
    return outs


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_noalloc_matches_submit_twice_second_smaller(ctx, oracle, flags):
    gops = [wl.mpeg2_gop(seed=41, gop_index=1, width=720, height=480, gop="IPBBPBB", field_fraction=0.3),
            wl.mpeg2_gop(seed=42, gop_index=2, width=176, height=144, gop="IPB")]
    results = []
    for units in gops:   # the larger GOP first, then the smaller one, back to back
        ref = _units_on_device(ctx, units)
        ctx.wait(ctx.submit_mpeg2([d.struct for d in ref], flags))
        got = _units_on_device(ctx, units)
        print(f"noalloc: {len(units)} pictures {units[0].meta['width']}x{units[0].meta['height']} flags {flags}", flush=True)
        ctx.reconstruct_mpeg2_noalloc([d.struct for d in got], flags)
        results.append((units, ref, got))
    for units, ref, got in results:
        want = _oracle(oracle, units)
        for i in range(len(units)):
            for n in ("y", "cb", "cr"):
                a, b = got[i].download(n), ref[i].download(n)
                assert np.array_equal(a, b), f"picture {i} {n}: allocation-free differs from mmr_recon_mpeg2_submit"
                assert np.array_equal(a, want[i][n]), f"picture {i} {n}: differs from the oracle"
        for d in ref + got:
            d.free()


def test_noalloc_rejects_more_than_its_scratch(ctx):
    """120 1080p pictures need 120 x 8160 x 96 B of block order (94 MB) > 64 MiB: rejected on the
    host before any upload or launch (the descriptors' buffers are small and never read; the same
    call on the host model of the runtime runs no kernel, tests/served_replay/checks.py noalloc)."""
    u = wl.mpeg2_gop(seed=43, gop_index=3, width=64, height=32, gop="I")[0]
    d = wl.DeviceUnit(ctx, u)
    big = abi.new(abi.Mpeg2Picture, width=1920, height=1080, picture_type=abi.MPEG2_PICTURE_I, matrices=d.ptr("matrices"),
                  macroblocks=d.ptr("mbs"), levels=d.ptr("levels"), forward_source=-1, backward_source=-1)
    for k, n in enumerate(("y", "cb", "cr")):
        big.planes[k] = d.ptr(n)
    with pytest.raises(mmrecode_amd.ReconError) as e:
        ctx.reconstruct_mpeg2_noalloc([big] * 120)
    assert e.value.status == abi.STATUS_INVALID_ARGUMENT
    assert "MMR_RECON_NOALLOC_SCRATCH_BYTES" in e.value.message
    # the context still works
    ctx.reconstruct_mpeg2_noalloc([d.struct])
    d.free()
