"""Write tests/golden/oracle_sha256.json: SHA-256 of seeded inputs and of the oracle's outputs.

These are SELF-GENERATED regression pins (the reference cannot be run here): they fix the
workload generators and the oracle's outputs on small cases so that any change to either is
detected. Regenerate only after a deliberate, reviewed change:  python tests/golden/make_golden.py
"""
import ctypes as C
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mmrecode_amd import abi  # noqa: E402
from mmrecode_amd import workloads as wl  # noqa: E402
from oracle import binding  # noqa: E402


def digest(arrays):
    h = hashlib.sha256()
    for name in sorted(arrays):
        h.update(name.encode())
        h.update(arrays[name].tobytes())
    return h.hexdigest()


def cases():
    L = binding.lib()
    out = {}
    for w, h, samp in ((640, 480, ((2, 2), (1, 1), (1, 1))), (33, 17, ((2, 1), (1, 1), (1, 1))), (21, 11, ((1, 1),))):
        u = wl.mjpeg_frame(7, w, w, h, samp)
        o = u.host_outputs()
        assert L.oracle_mjpeg_reconstruct(C.byref(u.host_struct(o))) == 0
        out[f"mjpeg_{w}x{h}_{len(samp)}c"] = {"inputs": digest(u.inputs), "outputs": digest(o)}
    units = wl.mpeg2_gop(7, 0, 176, 144)
    outs = {}
    for i, u in enumerate(units):
        outs[i] = u.host_outputs()
        assert L.oracle_mpeg2_reconstruct(C.byref(u.host_struct(outs[i], wl.mpeg2_ref_ptrs(outs, u, "host")))) == 0
    out["mpeg2_176x144_gop15"] = {"inputs": digest({f"{i}_{k}": v for i, u in enumerate(units) for k, v in u.inputs.items()}),
                                  "outputs": digest({f"{i}_{k}": v for i, o in outs.items() for k, v in o.items()})}
    for system in (abi.DV_SYSTEM_525_60, abi.DV_SYSTEM_625_50):
        u = wl.dv_frame(7, system, system)
        o = u.host_outputs()
        assert L.oracle_dv_reconstruct(C.byref(u.host_struct(o))) == 0
        out[f"dv_system{system}"] = {"inputs": digest(u.inputs), "outputs": digest(o)}
    refs = [dict(zip(("y", "cb", "cr"), wl.h264_reference_planes(7, 5, 4, i))) for i in range(4)]
    for deblock in (0, 1):
        u = wl.h264_picture(7, deblock, 5, 4, deblock=bool(deblock))
        o = u.host_outputs()
        assert L.oracle_h264_reconstruct(C.byref(u.host_struct(o, wl.h264_ref_ptrs(u, refs, "host")))) == 0
        out[f"h264_80x64_deblock{deblock}"] = {"inputs": digest({**u.inputs, **{f"r{i}{k}": v for i, r in enumerate(refs) for k, v in r.items()}}),
                                               "outputs": digest(o)}
    return out


if __name__ == "__main__":
    data = cases()
    with open(os.path.join(HERE, "oracle_sha256.json"), "w") as f:
        json.dump({"note": "self-generated regression pins, see make_golden.py", "cases": data}, f, indent=1, sort_keys=True)
    print(f"wrote {len(data)} cases")
