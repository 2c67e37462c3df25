"""The resident frame cache holding decoded MPEG-2 pictures: each picture of a GOP is reconstructed
into planes the cache creates (one entry per plane), its references are fetched from the cache,
and the byte budget holds four pictures, so older pictures are evicted least recently used first
while the anchors the next pictures need stay resident because fetching them touches them. Every
picture matches the oracle bit for bit, and a cache hit returns the same storage with the frame
still in it."""
import os
import sys

import pytest

from mmrecode_amd import DeviceBuffer, FrameCache
from mmrecode_amd import workloads as wl

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_parity import _assert_same, _mpeg2_oracle  # noqa: E402

pytestmark = pytest.mark.gpu

PLANES = ("y", "cb", "cr")


def test_decoded_pictures_resident_in_the_frame_cache(ctx, oracle):
    units = wl.mpeg2_gop(seed=31, gop_index=2, width=352, height=288, gop="IPBBPBBPBB")
    want = _mpeg2_oracle(oracle, units)
    shapes = {n: units[0].outputs[n][0] for n in PLANES}
    sizes = {n: shapes[n][0] * shapes[n][1] for n in PLANES}
    cache = FrameCache(ctx, 4 * sum(sizes.values()))

    def key(i, k):
        n = PLANES[k]
        return dict(namespace_id=1, owner=42, revision=1, local_frame=i, width=shapes[n][1], height=shapes[n][0], variant=k)

    def desc(k):
        n = PLANES[k]
        return dict(width=shapes[n][1], height=shapes[n][0], format=k + 1, color=1, alpha_mode=0)

    storage = {}
    for i, u in enumerate(units):
        cache.begin_frame()
        extra = {"forward_source": -1, "backward_source": -1}
        for tag in ("fwd", "bwd"):
            ref = u.meta[tag]
            if ref is not None:
                for k, n in enumerate(PLANES):
                    p = cache.get(key(ref, k), desc(k))
                    assert p is not None, (i, tag, ref)
                    extra[f"{tag}_{n}"] = p
        for k, n in enumerate(PLANES):
            p, _ = cache.retain(key(i, k), desc(k), sizes[n])
            extra[n] = p
            storage[(i, k)] = p
        d = wl.DeviceUnit(ctx, u, extra)
        ctx.wait(ctx.submit_mpeg2([d.struct]))
        for k, n in enumerate(PLANES):
            got = ctx.download(DeviceBuffer(ctx, extra[n], sizes[n])).reshape(shapes[n])
            _assert_same(f"picture {i} ({u.meta['type']}) {n}", got, want[i][n])
        d.free()
    st = cache.stats()
    assert st["evictions"] > 0 and st["retained_bytes"] <= 4 * sum(sizes.values())
    # the last picture is still resident: a hit gives its storage back with the picture in it
    last = len(units) - 1
    p, status = cache.retain(key(last, 0), desc(0), sizes["y"])
    assert p == storage[(last, 0)] and status == 2
    _assert_same("resident last picture", ctx.download(DeviceBuffer(ctx, p, sizes["y"])).reshape(shapes["y"]), want[last]["y"])
    cache.close()
