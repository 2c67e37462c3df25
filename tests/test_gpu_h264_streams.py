"""H.264 bitstreams on the device: CAVLC streams decoded by the oracle's front end
(oracle/h264_entropy.c) into records, reconstructed picture by picture in decoding order on the
GPU with each picture's references resident on the device, and compared bit for bit with the
oracle's reconstruction of the same records; the reference's weighted bi-prediction scenes
(crates/codecs/h264/src/decoder.rs:16853-16950) also give their known planes on the device."""
import os
import sys

import numpy as np
import pytest

from mmrecode_amd import workloads as wl

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import h264_streams as hs  # noqa: E402

pytestmark = pytest.mark.gpu


def reconstruct_on_device(ctx, units):
    planes, dev = {}, []
    for u in units:
        d = wl.DeviceUnit(ctx, u, hs.h264_refs(u, planes, "dev"))
        ctx.wait(ctx.submit_h264([d.struct]))
        planes[u.meta["slot"]] = {n: d.ptr(n) for n in ("y", "cb", "cr")}
        dev.append(d)
    return dev


@pytest.mark.parametrize("mode, want", [("explicit", (88, 105, 150)), ("implicit", (57, 95, 150))])
def test_reference_weighted_bipred_streams_on_device(ctx, oracle, mode, want):
    units = [hs.unit_of(d) for d in hs.decode_stream(oracle, hs.weighted_bipred_stream(mode))]
    dev = reconstruct_on_device(ctx, units)
    for n, v in zip(("y", "cb", "cr"), want):
        got = dev[-1].download(n)
        assert (got == v).all(), (mode, n, np.unique(got))
    for d in dev:
        d.free()


@pytest.mark.parametrize("seed, w, h, pattern, high", [(1, 5, 4, "IPBPB", True), (2, 3, 3, "IPPBB", False),
                                                       (3, 7, 2, "IPBPBB", True), (5, 4, 5, "IBPB", True),
                                                       (11, 12, 9, "IPBBPB", True)])
def test_cavlc_streams_bit_exact(ctx, oracle, seed, w, h, pattern, high):
    aus, _ = hs.random_stream(seed, w, h, pattern, high=high)
    units = [hs.unit_of(d) for d in hs.decode_stream(oracle, aus)]
    want = hs.reconstruct_on_oracle(oracle, units)
    dev = reconstruct_on_device(ctx, units)
    for i, (u, d) in enumerate(zip(units, dev)):
        for n in ("y", "cb", "cr"):
            got, ref = d.download(n), want[u.meta["slot"]][n]
            assert np.array_equal(got, ref), f"picture {i} {n}: {int((got != ref).sum())} samples differ"
    for d in dev:
        d.free()
