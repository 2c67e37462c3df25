"""The reference's three MPEG-2 test streams and the video of its two transport-stream vectors,
entropy-decoded by the oracle's front end and reconstructed on the device -- progressive,
interlaced with field prediction, and open GOPs with cross-GOP references -- bit-exact against
the oracle for every picture, as one submission (dependency-level launches) and with one launch per
picture."""
import pytest

import mpeg2_streams as ms
from mmrecode_amd import abi
from test_gpu_parity import _assert_same, _mpeg2_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
@pytest.mark.parametrize("name", sorted(ms.STREAMS) + sorted(ms.TS_STREAMS))
def test_mpeg2_stream_bit_exact(ctx, oracle, name, flags):
    from oracle import binding

    pics = binding.mpeg2_pictures(ms.read(name))
    units = ms.units(pics)
    want = ms.oracle_decode(oracle, units)
    dev = _mpeg2_gpu(ctx, units, flags)
    try:
        for i, d in enumerate(dev):
            for n in ("y", "cb", "cr"):
                _assert_same(f"{name} picture {i} (display {pics[i]['display_index']}) {n}", d.download(n), want[i][n])
    finally:
        for d in dev:
            d.free()
