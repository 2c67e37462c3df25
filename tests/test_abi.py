"""The C ABI boundary: header hygiene, struct layouts shared with the ctypes mirror, exports.

Mirrors crates/capi/tests/header.rs:6-24 (compile the header as C11) and adds layout and export
checks. No compute call is made, so these run without a GPU.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from mmrecode_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mmr_recon.h")
LIB = os.path.join(ROOT, "mmrecode_amd", "libmmr_recon.so")

STRUCTS = {
    "mmr_recon_config": abi.Config,
    "mmr_recon_mjpeg_frame": abi.MjpegFrame,
    "mmr_recon_mpeg2_macroblock": abi.Mpeg2Macroblock,
    "mmr_recon_mpeg2_picture": abi.Mpeg2Picture,
    "mmr_recon_mpeg2_field_motion": abi.Mpeg2FieldMotion,
    "mmr_recon_dv_frame": abi.DvFrame,
    "mmr_recon_h264_macroblock": abi.H264Macroblock,
    "mmr_recon_h264_block": abi.H264Block,
    "mmr_recon_h264_intra": abi.H264Intra,
    "mmr_recon_h264_weight": abi.H264Weight,
    "mmr_recon_h264_bi_weight": abi.H264BiWeight,
    "mmr_recon_h264_reference": abi.H264Reference,
    "mmr_recon_h264_picture": abi.H264Picture,
    "mmr_recon_yuv_to_rgb_job": abi.YuvToRgbJob,
    "mmr_recon_layer": abi.Layer,
    "mmr_recon_composite_job": abi.CompositeJob,
    "mmr_recon_linear_layer": abi.LinearLayer,
    "mmr_recon_linear_composite_job": abi.LinearCompositeJob,
    "mmr_recon_linear_convert_job": abi.LinearConvertJob,
    "mmr_recon_frame_key": abi.FrameKey,
    "mmr_recon_frame_descriptor": abi.FrameDescriptor,
    "mmr_recon_frame_cache_stats": abi.FrameCacheStats,
    "mmr_recon_yuv_blend_job": abi.YuvBlendJob,
    "mmr_recon_overlay_prepare_job": abi.OverlayPrepareJob,
    "mmr_recon_scale_job": abi.ScaleJob,
    "mmr_recon_server_config": abi.ServerConfig,
    "mmr_recon_request": abi.Request,
    "mmr_recon_frame_completion": abi.FrameCompletion,
    "mmr_recon_server_stats": abi.ServerStats,
    "mmr_recon_tuning": abi.Tuning,
    "mmr_recon_tuning_trial": abi.TuningTrial,
}

needs_cc = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not available")


@needs_cc
def test_header_compiles_as_strict_c11(tmp_path):
    src = tmp_path / "h.c"
    src.write_text('#include "mmr_recon.h"\nint main(void) { return (int)MMR_RECON_ABI_VERSION - 1; }\n')
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.dirname(HEADER),
                    str(src), "-o", str(tmp_path / "h")], check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++", "-fsyntax-only", "-I",
                    os.path.dirname(HEADER), str(src)], check=True)


@needs_cc
def test_struct_layouts_match_ctypes(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mmr_recon.h"', 'int main(void) {']
    for cname, ctype in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in ctype._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, value = line.split()
        ctype = STRUCTS[cname]
        got = C.sizeof(ctype) if field == "size" else getattr(ctype, field).offset
        assert got == int(value), f"{cname}.{field}: ctypes {got} != C {value}"


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"MMR_RECON_API\s+[\w\s\*]+?\b(mmr_recon_\w+)\s*\(", text)))


def test_every_declared_function_is_exported_and_bound():
    assert os.path.exists(LIB), "libmmr_recon.so is not built (python -m mmrecode_amd.build)"
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 20
    bound = {name for name, _, _ in abi.ENTRY_POINTS}
    for name in names:
        assert hasattr(lib, name), f"{name} is declared in mmr_recon.h but not exported"
        assert name in bound, f"{name} has no ctypes signature in mmrecode_amd.abi"
    assert lib.mmr_recon_abi_version() == abi.ABI_VERSION


def test_exports_are_only_the_c_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert exported, "no exported text symbols"
    stray = sorted(s for s in exported if not s.startswith("mmr_recon_") and not s.startswith("_"))
    assert not stray, f"unexpected exports: {stray[:10]}"


def test_host_only_entry_points_without_gpu():
    """Calls that touch no device: version, error message, argument validation, implicit weights."""
    import mmrecode_amd

    lib = mmrecode_amd.load_library()
    assert lib.mmr_recon_abi_version() == 1
    w0, w1 = C.c_int32(), C.c_int32()
    assert lib.mmr_recon_h264_implicit_weights(0, 4, 1, 0, 0, C.byref(w0), C.byref(w1)) == abi.STATUS_OK
    assert (w0.value, w1.value) == (48, 16)
    assert lib.mmr_recon_h264_implicit_weights(0, 4, 1, 0, 0, None, None) == abi.STATUS_INVALID_ARGUMENT
    assert "NULL" in mmrecode_amd.last_error()
    assert lib.mmr_recon_context_create(None, None) == abi.STATUS_INVALID_ARGUMENT
    bad = abi.new(abi.Config)
    bad.struct_size = 4
    handle = C.c_void_p()
    assert lib.mmr_recon_context_create(C.byref(bad), C.byref(handle)) == abi.STATUS_INVALID_ARGUMENT
    assert lib.mmr_recon_wait(None, 1) == abi.STATUS_INVALID_ARGUMENT
    n = lib.mmr_recon_last_error_message(None, 0)
    assert n == len(mmrecode_amd.last_error()) + 1
