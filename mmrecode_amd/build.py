"""Build libmmr_recon.so (HIP, gfx950) in-tree with hipcc.

Every HIP source is compiled with -ffp-contract=off (the f64 IDCTs must not fuse a*b+c, as the
reference never does) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmmr_recon.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["recon_host.cpp", "mjpeg.hip", "mpeg2.hip", "dv.hip", "h264.hip", "h264_intra.hip", "deblock.hip", "color.hip"]
ARCH = os.environ.get("MMR_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-fvisibility=hidden",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    # The .cpp host file holds HIP runtime calls only; every file is compiled as HIP.
    cmd = [HIPCC, *FLAGS, "-o", tmp, "-x", "hip", *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
