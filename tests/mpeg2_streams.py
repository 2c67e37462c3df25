"""Shared helpers for the MPEG-2 elementary-stream tests: the reference's three test streams
(testdata/mpeg2/valid, kept under tests/golden/mpeg2), entropy-decoded by the oracle's front end
into reconstruction units, and the oracle's reconstruction of every picture."""
import ctypes as C
import os

from mmrecode_amd import workloads as wl

HERE = os.path.join(os.path.dirname(__file__), "golden", "mpeg2")

# name -> pictures in the stream (crates/codecs/mpeg2/tests/reference_vectors.rs:19-20, :61, :92)
STREAMS = {
    "main-ml-progressive-ibp.m2v": 12,
    "main-ml-interlaced-ibp.m2v": 12,
    "main-ml-progressive-open-gop.m2v": 24,
}


def read(name):
    with open(os.path.join(HERE, name), "rb") as f:
        return f.read()


def units(pictures):
    """Reconstruction units of the pictures mpeg2_pictures returned, in decode order."""
    out = []
    for p in pictures:
        field = p["field_motion"] if p["field_motion_mbs"] else None
        out.append(wl.mpeg2_unit(p["width"], p["height"], p["type"], p["matrices"], p["mbs"], p["levels"], field,
                                 p["alternate_scan"], p["intra_dc_precision"], p["fwd"], p["bwd"],
                                 {"type": p["type"], "field_mbs": p["field_motion_mbs"]}))
    return out


def oracle_decode(oracle, us):
    """The oracle's reconstruction of every unit, references resolved through earlier outputs."""
    outs = {}
    for i, u in enumerate(us):
        out = u.host_outputs()
        s = u.host_struct(out, wl.mpeg2_ref_ptrs(outs, u, "host"))
        rc = oracle.oracle_mpeg2_reconstruct(C.byref(s))
        if rc != 0:
            raise AssertionError(f"oracle_mpeg2_reconstruct picture {i}: {rc}")
        outs[i] = out
    return outs
