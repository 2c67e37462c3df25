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


TS_HERE = os.path.join(os.path.dirname(__file__), "golden", "mpegts")

# MPEG-2 transport streams (testdata/mpegts/valid): name -> pictures of the video elementary stream
# on PID 0x0100 (crates/containers/mpegts/tests/reference_vectors.rs:20-38)
TS_STREAMS = {
    "single-program-mpeg2.ts": 12,
    "single-program-mpeg2-mp2.ts": 12,
}
TS_VIDEO_PID = 0x0100


def read(name):
    """The elementary stream of a test vector: an .m2v as it is, a transport stream demuxed."""
    if name in TS_STREAMS:
        with open(os.path.join(TS_HERE, name), "rb") as f:
            return ts_video_es(f.read())
    with open(os.path.join(HERE, name), "rb") as f:
        return f.read()


def ts_video_es(data, pid=TS_VIDEO_PID):
    """The payload bytes of every PES packet on `pid` of an ISO/IEC 13818-1 transport stream, in
    order: 188-byte packets with sync byte 0x47, the adaptation field skipped, continuity counters
    checked, PES packets started by payload_unit_start_indicator, their headers
    (9 + PES_header_data_length bytes) dropped. Raises ValueError on damage. Test plumbing for the
    reference's transport-stream vectors, not a demultiplexer."""
    if len(data) % 188:
        raise ValueError("transport stream length is not a multiple of 188")
    pes, packets, cc = None, [], None
    for off in range(0, len(data), 188):
        p = data[off:off + 188]
        if p[0] != 0x47:
            raise ValueError(f"lost sync at byte {off}")
        if (((p[1] & 0x1F) << 8) | p[2]) != pid:
            continue
        afc, counter = (p[3] >> 4) & 3, p[3] & 0x0F
        if afc & 1:
            if cc is not None and counter != (cc + 1) & 0x0F:
                raise ValueError(f"continuity counter {counter} after {cc} at byte {off}")
            cc = counter
        start = 4 + (1 + p[4] if afc & 2 else 0)
        if not afc & 1:
            continue
        if (p[1] >> 6) & 1:
            if pes is not None:
                packets.append(pes)
            pes = bytearray(p[start:])
        elif pes is not None:
            pes += p[start:]
    if pes is not None:
        packets.append(pes)
    es = bytearray()
    for q in packets:
        if q[:3] != b"\x00\x00\x01" or len(q) < 9 or len(q) < 9 + q[8]:
            raise ValueError("PES packet without a start code or with a truncated header")
        es += q[9 + q[8]:]
    return bytes(es)


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
