"""CPU replay of the served mixed workload through the library's real host code.

Loads libmmr_recon_model.so (the library's recon_host.cpp on the host model of the HIP runtime,
with address models of the kernels; see build.py) in place of libmmr_recon.so and drives it with
bench.py's own ServedMix: the same pools, 64 streams, ring slots, request windows and descriptors
that `bench.py --served` submits, through the same continuous-batching server (mmr_recon_server_*),
then the same trace through the plain one-frame-per-launch path. Every buffer range a kernel would
touch is checked against the allocations live when the kernel runs (stream order, with streams
interleaved at random); the replay reports the violations it found.

python tests/served_replay/replay.py [--rates 50,2000] [--seconds 1.0] [--plain-seconds 5] [--seed 7]
Prints one JSON line; exit status 1 when the model reports a violation or a frame fails.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import build as model_build  # noqa: E402


def load_model_library():
    """Builds the model library and makes mmrecode_amd load it instead of libmmr_recon.so (call
    before the first Context of the process)."""
    import mmrecode_amd

    path = model_build.build()
    if mmrecode_amd._LIB is not None and mmrecode_amd.LIBRARY_PATH != path:
        raise RuntimeError("libmmr_recon.so is already loaded in this process")
    mmrecode_amd.LIBRARY_PATH = path
    lib = mmrecode_amd.load_library()
    lib.mmr_model_violations.restype = C.c_size_t
    lib.mmr_model_violations.argtypes = [C.c_char_p, C.c_size_t]
    lib.mmr_model_stats.restype = None
    lib.mmr_model_stats.argtypes = [C.POINTER(C.c_uint64)]
    lib.mmr_model_seed.restype = None
    lib.mmr_model_seed.argtypes = [C.c_uint64]
    return lib


def violations(lib):
    buf = C.create_string_buffer(1 << 16)
    n = lib.mmr_model_violations(buf, len(buf))
    return n, [v for v in buf.value.decode(errors="replace").split("\n") if v]


def model_stats(lib):
    out = (C.c_uint64 * 6)()
    lib.mmr_model_stats(out)
    return {"kernels": out[0], "checks": out[1], "ops": out[2], "jobs": out[3], "allocations": out[4], "events": out[5]}


def replay(rates, seconds, plain_seconds, seed=7, slots=4, log=print):
    lib = load_model_library()
    lib.mmr_model_seed(seed)
    import numpy as np

    import bench
    import mmrecode_amd

    ctx = mmrecode_amd.Context(device=0)
    t = time.time()
    mix = bench.ServedMix(ctx)
    log(f"replay: pools built in {time.time() - t:.1f}s")
    rng = np.random.default_rng(4242)
    runs = []
    for rate in rates:
        reqs = mix.build(mix.trace(rate, seconds, rng))
        served = mix.run_served(reqs, slots=slots)
        entry = {"rate_req_s": rate, "requests": len(reqs), "frames": sum(r["n"] for r in reqs),
                 "served_frames": served["frames"], "bad_status": served["bad_status"], "batches": served["batches"]}
        if plain_seconds > 0:
            plain = mix.run_plain(reqs, plain_seconds)
            entry["plain_frames"] = plain["frames"]
        n, _ = violations(lib)
        entry["violations_so_far"] = n
        log(f"replay: {entry}")
        runs.append(entry)
    ctx.close()
    n, msgs = violations(lib)
    return {"runs": runs, "violations": n, "messages": msgs[:20], "model": model_stats(lib)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="50,500,2000")
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--plain-seconds", type=float, default=3.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--slots", type=int, default=4)
    args = ap.parse_args()
    out = replay([float(r) for r in args.rates.split(",")], args.seconds, args.plain_seconds, args.seed, args.slots,
                 log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out))
    bad = out["violations"] or any(r["bad_status"] or r["served_frames"] != r["frames"] for r in out["runs"])
    sys.exit(1 if bad else 0)


if __name__ == "__main__":
    main()
