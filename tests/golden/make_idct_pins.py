"""Generates tests/golden/idct_pins.json: coefficient blocks on which the rounding of the reference's
f64 IDCTs is decided by IEEE-754 details, for tests/test_oracle_idct_pins.py.

For each transform (MJPEG and MPEG-2 inverse_dct, DV inverse_88 and inverse_248; loop orders and
sources in tests/idct_reference.py) a seeded search over sparse blocks keeps blocks where
  * "tie": the reference's binary64 evaluation lands exactly on k + 1/2, so round half away from
    zero (Rust's f64::round) and round half to even give different samples, and
  * "fused": evaluating the same order with each a * b + s rounded once (FMA contraction) rounds a
    sample to a different integer than the reference's separate product and sum roundings.
"mpeg2_odd" holds MPEG-2 blocks of both kinds whose coefficient sum is odd (left unchanged by
mismatch control). The reference evaluation is tests/idct_reference.py's `plain` (Python binary64
floats, never fused);
the fused one uses exact rationals with one rounding per fused step.

DV inverse_88 sums 64 four-factor products per sample, so blocks on which fusion changes a rounded
sample are rare: the two-minute sparse search finds none. Its fused entries come from
`search_dv88_fused` (seeded, 3 to 11 coefficients up to +-300), which found the block in
DV88_FUSED_FOUND after 1,135,171 blocks (about 20 minutes); main() checks that block again and keeps
it. `--dv88-search SECONDS` runs that search anew.

python tests/golden/make_idct_pins.py [--dv88-search SECONDS]   (a few minutes)
"""
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from idct_reference import Transforms, round_half_away  # noqa: E402


def block(rng, kind):
    c = [0] * 64
    k = int(rng.integers(2, 5))
    for i in rng.choice(64, k, replace=False):
        c[int(i)] = int(rng.integers(-12, 13))
    if kind.startswith("dv"):
        c = [v * int(rng.choice([1, 2, 4, 8])) for v in c]
        c[0] = int(rng.integers(0, 2048))
    elif kind == "mjpeg" and rng.random() < 0.5:
        c[0] = int(rng.integers(-1024, 1024))
    return c


def search(T, kind, seconds, want=3, seed=2, odd=False):
    rng = np.random.default_rng([seed, T.KINDS.index(kind), int(odd)])
    ties, fuse, tried = [], [], 0
    t = time.time()
    while time.time() - t < seconds and (len(ties) < want or len(fuse) < want):
        c = block(rng, kind)
        if odd and sum(c) % 2 == 0:
            continue
        tried += 1
        ev = T.plain(kind, c)
        near = [i for i, v in enumerate(ev) if abs(abs(v - math.floor(v)) - 0.5) < 1e-9]
        if not near:
            continue
        if len(ties) < want and any(ev[i] - math.floor(ev[i]) == 0.5 for i in near):
            ties.append(c)
        if len(fuse) < want:
            fv = T.fused(kind, c)
            if any(round_half_away(fv[i]) != round_half_away(ev[i]) for i in near):
                fuse.append(c)
    return ties, fuse, tried


DV88_FUSED_FOUND = [{"0": 289, "4": -105, "32": 110}]


def fused_differs(T, kind, c):
    plain, fused = T.plain(kind, c), T.fused(kind, c)
    return any(round_half_away(a) != round_half_away(b) for a, b in zip(plain, fused))


def search_dv88_fused(T, seconds, want=3, seed=77):
    rng = np.random.default_rng(seed)
    found, tried = [], 0
    t = time.time()
    while time.time() - t < seconds and len(found) < want:
        c = [0] * 64
        for i in rng.choice(64, int(rng.integers(3, 12)), replace=False):
            c[int(i)] = int(rng.integers(-300, 301))
        c[0] = int(rng.integers(0, 2048))
        tried += 1
        ev = T.plain("dv88", c)
        if any(abs(abs(v - math.floor(v)) - 0.5) < 1e-11 for v in ev) and fused_differs(T, "dv88", c):
            found.append({str(i): v for i, v in enumerate(c) if v})
            print(f"dv88 fused block after {tried} blocks: {found[-1]}", file=sys.stderr, flush=True)
    return found


def main():
    from oracle import binding

    T = Transforms(binding.lib())
    dv88_seconds = float(sys.argv[sys.argv.index("--dv88-search") + 1]) if "--dv88-search" in sys.argv else 0
    out = {}
    for kind in T.KINDS:
        ties, fuse, tried = search(T, kind, 120)
        print(f"{kind}: {tried} blocks tried, {len(ties)} tie blocks, {len(fuse)} fused blocks", file=sys.stderr, flush=True)
        out[kind] = {"tie": [{str(i): v for i, v in enumerate(c) if v} for c in ties],
                     "fused": [{str(i): v for i, v in enumerate(c) if v} for c in fuse]}
    extra = search_dv88_fused(T, dv88_seconds) if dv88_seconds else []
    for d in DV88_FUSED_FOUND + extra:
        c = [0] * 64
        for i, v in d.items():
            c[int(i)] = v
        assert fused_differs(T, "dv88", c), d
        if d not in out["dv88"]["fused"]:
            out["dv88"]["fused"].append(d)
    # MPEG-2 blocks whose coefficient sum is odd: mismatch control (H.262 7.4.4) leaves them as they
    # are, so dequantised levels reproduce them exactly on the device (tests/test_gpu_idct_fast_path.py)
    ties, fuse, tried = search(T, "mpeg2", 240, want=4, odd=True)
    print(f"mpeg2 odd-sum: {tried} blocks tried, {len(ties)} tie blocks, {len(fuse)} fused blocks", file=sys.stderr, flush=True)
    out["mpeg2_odd"] = {"tie": [{str(i): v for i, v in enumerate(c) if v} for c in ties],
                        "fused": [{str(i): v for i, v in enumerate(c) if v} for c in fuse]}
    with open(os.path.join(HERE, "idct_pins.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
