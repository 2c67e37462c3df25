"""Real JPEG / MJPEG streams (the reference's corpus, entropy-decoded on the CPU) reconstructed on
the device through the C ABI, bit-exact against the oracle; the restart-interval vector is the
reference's flat-128 known answer."""
import numpy as np
import pytest

from oracle import binding
from mmrecode_amd import workloads as wl

from test_gpu_parity import _assert_same, _oracle_run
from test_jpeg_streams import VECTORS, read

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", VECTORS)
def test_jpeg_vector_bit_exact(ctx, oracle, name):
    frames = binding.jpeg_frames(read(name))
    units = [wl.mjpeg_unit_from_frame(f) for f in frames]
    dev = [wl.DeviceUnit(ctx, u) for u in units]
    ctx.wait(ctx.submit_mjpeg([d.struct for d in dev]))   # all frames of the file in one launch
    for i, (u, d) in enumerate(zip(units, dev)):
        want = _oracle_run(oracle, "oracle_mjpeg_reconstruct", u)
        for n in want:
            _assert_same(f"{name} frame {i} {n}", d.download(n), want[n])
        if name == "restart-interval.jpg":
            assert np.all(d.download("plane0") == 128)
        d.free()
