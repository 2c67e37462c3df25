"""Real JPEG / MJPEG streams of the reference's corpus through the CPU entropy front end (test
infrastructure, oracle/jpeg_entropy.c) and the oracle's reconstruction: the front end's known
answers, the reference decoder's assertions on these files, and a known-answer picture."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

from oracle import binding
from mmrecode_amd import workloads as wl

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jpeg")
VECTORS = ["baseline-420.jpg", "minimal-gray.jpg", "unknown-app-marker.jpg", "restart-interval.jpg", "mmrecode-q85-420.mjpg"]


def read(name):
    with open(os.path.join(HERE, name), "rb") as f:
        return f.read()


def reconstruct(oracle, frame, rgb=True):
    u = wl.mjpeg_unit_from_frame(frame, rgb=rgb)
    out = u.host_outputs()
    s = u.host_struct(out)
    assert oracle.oracle_mjpeg_reconstruct(C.byref(s)) == 0
    return out


def test_vectors_are_the_reference_corpus():
    sums = dict(line.split()[::-1] for line in open(os.path.join(HERE, "SHA256SUMS")).read().splitlines())
    for name in VECTORS:
        assert hashlib.sha256(read(name)).hexdigest() == sums[name]


def test_canonical_codes_stuffing_and_extension(oracle):
    # entropy.rs:190-210: table {1, 2} codes of lengths 1, 2, 2 for symbols 10, 20, 30
    counts = np.zeros(16, np.uint8)
    counts[:2] = [1, 2]
    symbols = np.array([10, 20, 30], np.uint8)
    data = np.array([0b0101_1000], np.uint8)
    out = np.zeros(3, np.int32)
    assert oracle.oracle_jpeg_huffman_symbols(counts.ctypes.data, symbols.ctypes.data, 3, data.ctypes.data, 1,
                                              out.ctypes.data, 3) == 0
    assert out.tolist() == [10, 20, 30]
    stuffed = np.array([0xFF, 0x00], np.uint8)
    v = C.c_uint32()
    assert oracle.oracle_jpeg_read_bits(stuffed.ctypes.data, 2, 8, C.byref(v)) == 0 and v.value == 0xFF
    bits = np.array([0b0011_0000], np.uint8)
    sizes = np.array([3, 3], np.int32)
    vals = np.zeros(2, np.int32)
    assert oracle.oracle_jpeg_receive_extend(bits.ctypes.data, 1, sizes.ctypes.data, 2, vals.ctypes.data) == 0
    assert vals.tolist() == [-6, 4]
    # an oversubscribed table is rejected
    over = np.zeros(16, np.uint8)
    over[0] = 3
    assert oracle.oracle_jpeg_huffman_symbols(over.ctypes.data, symbols.ctypes.data, 3, data.ctypes.data, 1,
                                              out.ctypes.data, 1) != 0


def test_reference_decoder_assertions(oracle):
    # decoder.rs:583-609: dimensions, component counts and plane sizes
    for name, dims, ncomp in [("baseline-420.jpg", (16, 16), 3), ("minimal-gray.jpg", (8, 8), 3),
                              ("unknown-app-marker.jpg", (16, 16), 3), ("restart-interval.jpg", (16, 8), 1)]:
        (f,) = binding.jpeg_frames(read(name))
        assert (f["width"], f["height"]) == dims and len(f["sampling"]) == ncomp
    (f,) = binding.jpeg_frames(read("baseline-420.jpg"))
    out = reconstruct(oracle, f)
    assert [out[f"plane{c}"].shape[::-1] for c in range(3)] == [(16, 16), (8, 8), (8, 8)]


def test_restart_interval_vector_is_flat_128(oracle):
    # decoder.rs:611-618: every reconstructed sample of restart-interval.jpg is 128
    (f,) = binding.jpeg_frames(read("restart-interval.jpg"))
    out = reconstruct(oracle, f)
    assert out["plane0"].shape == (8, 16) and np.all(out["plane0"] == 128)


def test_app_marker_does_not_change_the_image(oracle):
    # testdata/README.md: unknown-app-marker.jpg is baseline-420.jpg plus an APP2 segment
    (a,) = binding.jpeg_frames(read("baseline-420.jpg"))
    (b,) = binding.jpeg_frames(read("unknown-app-marker.jpg"))
    oa, ob = reconstruct(oracle, a), reconstruct(oracle, b)
    for n in oa:
        assert np.array_equal(oa[n], ob[n])


def test_mjpeg_file_holds_two_frames(oracle):
    frames = binding.jpeg_frames(read("mmrecode-q85-420.mjpg"))
    assert len(frames) == 2
    for f in frames:
        assert (f["width"], f["height"], f["sampling"]) == (16, 16, [(2, 2), (1, 1), (1, 1)])
        out = reconstruct(oracle, f)
        assert out["plane0"].std() > 0   # real picture content, not a flat fill


def test_truncated_stream_is_rejected():
    data = read("baseline-420.jpg")
    with pytest.raises(binding.JpegError):
        binding.jpeg_frames(data[: len(data) // 2])
    with pytest.raises(binding.JpegError):
        binding.jpeg_frames(b"\x00\x01" + data[2:])
