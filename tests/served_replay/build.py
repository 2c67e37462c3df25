"""Builds libmmr_recon_model.so: the library's real host code (mmrecode_amd/csrc/recon_host.cpp)
compiled with g++ against the host model of the HIP runtime (hip/hip_runtime.h, hip_model.cpp) and
the kernels' address models (kernel_model.cpp). Test infrastructure for the CPU served-path replay
and the exception-containment test; it never runs on a GPU and the product does not link it."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libmmr_recon_model.so")
SOURCES = [os.path.join(ROOT, "mmrecode_amd", "csrc", "recon_host.cpp"), os.path.join(HERE, "hip_model.cpp"),
           os.path.join(HERE, "kernel_model.cpp")]


def build(verbose=False):
    deps = SOURCES + [os.path.join(HERE, f) for f in ("model.h", os.path.join("hip", "hip_runtime.h"))] + [
        os.path.join(ROOT, "mmrecode_amd", "csrc", "recon_internal.h"), os.path.join(ROOT, "include", "mmr_recon.h")]
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(d) for d in deps):
        return LIB
    # a per-process temporary name: parallel test workers may build at the same time
    tmp = f"{LIB}.tmp{os.getpid()}"
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-g", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall",
           "-Wno-unused-function", "-I", HERE, "-o", tmp, *SOURCES, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
