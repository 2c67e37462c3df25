"""The reference's two DV25 test frames (525/60 4:1:1 and 625/50 4:2:0), entropy-decoded by the
oracle's front end and reconstructed on the device bit-exact against the oracle: both frames in one
submission and one launch per frame, and a frame whose first segment was concealed."""
import pytest

from mmrecode_amd import abi
from mmrecode_amd import workloads as wl

from test_dv_streams import FRAMES, decode, read
from test_gpu_parity import _assert_same

pytestmark = pytest.mark.gpu


def _units(conceal_first=False):
    from oracle import binding

    out = []
    for name in sorted(FRAMES):
        data = bytearray(read(name))
        if conceal_first:
            data[7 * 80] = 0x76
        f = binding.dv_frame_levels(bytes(data), conceal=conceal_first)
        out.append((name, f, wl.dv_unit(f["system"], f["levels"], f["info"])))
    return out


@pytest.mark.parametrize("conceal", [False, True])
@pytest.mark.parametrize("flags", [0, abi.SUBMIT_PER_FRAME_LAUNCH])
def test_dv_frames_bit_exact(ctx, oracle, flags, conceal):
    units = _units(conceal)
    dev = [wl.DeviceUnit(ctx, u) for _, _, u in units]
    try:
        ctx.wait(ctx.submit_dv([d.struct for d in dev], flags))
        for (name, f, _), d in zip(units, dev):
            want = decode(oracle, f)
            for n in ("y", "cb", "cr"):
                _assert_same(f"{name} {n}", d.download(n), want[n])
    finally:
        for d in dev:
            d.free()
