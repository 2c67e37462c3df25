"""Generates tests/golden/dv_fallback_blocks.json: DV blocks, as the device takes them (levels by
scan position and the info byte), on which the DV kernel's fused transform without its tie window
would round a sample differently from the reference's order (inverse_88 / inverse_248), found by
the seeded search of tests/fastpath/dv_fast_check.c. tests/test_idct_fast_path.py checks each again
on the CPU; tests/test_gpu_idct_fast_path.py reconstructs a frame holding them on the device.

python tests/golden/make_dv_fallback_blocks.py   (seconds)
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SEEDS = {"88": 11, "248": 12}
WANT = 4


def build(out_dir):
    exe = os.path.join(out_dir, "dv_fast_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", exe,
                    os.path.join(ROOT, "tests", "fastpath", "dv_fast_check.c"), "-ldl", "-lm"], check=True)
    return exe


def main():
    sys.path.insert(0, ROOT)
    from oracle import binding

    binding.lib()
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        exe = build(d)
        for mode, seed in SEEDS.items():
            r = subprocess.run([exe, so, "search", mode, "1000000", str(seed), str(WANT)], stdin=subprocess.DEVNULL,
                               capture_output=True, text=True, check=True, timeout=300)
            blocks = []
            for line in r.stdout.split("\n"):
                if not line.strip():
                    continue
                info, *pairs = line.split()
                blocks.append({"info": int(info), "levels": {p.split(":")[0]: int(p.split(":")[1]) for p in pairs}})
            assert len(blocks) == WANT, (mode, r.stdout)
            out["dv" + mode] = blocks
    with open(os.path.join(HERE, "dv_fallback_blocks.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
