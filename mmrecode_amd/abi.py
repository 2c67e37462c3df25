"""ctypes mirror of include/mmr_recon.h (struct layouts and constants).

The layouts must match the C header byte for byte; tests/test_abi.py checks the
sizes and field offsets against a C compiler's view of the header.
"""
import ctypes as C

ABI_VERSION = 1
MAX_PHASES = 4            # MMR_RECON_MAX_PHASES

STATUS_OK = 0
STATUS_INVALID_ARGUMENT = 1
STATUS_INVALID_DATA = 2
STATUS_UNSUPPORTED = 3
STATUS_INVALID_STATE = 4
STATUS_IO_ERROR = 5
STATUS_BUSY = 6
STATUS_INTERNAL_ERROR = 100

RANGE_FULL = 1
RANGE_LIMITED = 2
MATRIX_BT601 = 1
MATRIX_BT709 = 2

SUBMIT_PER_FRAME_LAUNCH = 1

MPEG2_MB_INTRA = 1
MPEG2_MB_FORWARD = 2
MPEG2_MB_BACKWARD = 4
MPEG2_MB_INTERLACED_DCT = 8
MPEG2_MB_FIELD_MOTION = 16
MPEG2_PICTURE_I = 1
MPEG2_PICTURE_P = 2
MPEG2_PICTURE_B = 3

DV_SYSTEM_525_60 = 0
DV_SYSTEM_625_50 = 1

H264_MB_INTRA = 1
H264_MB_TRANSFORM_8X8 = 2
H264_MB_NO_DEBLOCK = 4
H264_MB_SLICE_EDGE_OFF = 8
H264_MAX_REFS = 16
H264_LEVELS_PER_MB = 384
H264_INTRA_4X4 = 0
H264_INTRA_8X8 = 1
H264_INTRA_16X16 = 2
H264_INTRA_PCM = 3
H264_WEIGHTED_NONE = 0
H264_WEIGHTED_TABLE = 1

BLEND_SOURCE_OVER = 1
BLEND_PREMULTIPLIED_OVER = 2
MAX_LAYERS = 8
LINEAR_SOURCE_SURFACE = 1
LINEAR_SOURCE_SRGBA8 = 2
LINEAR_TO_SRGBA8 = 1
LINEAR_FROM_SRGBA8 = 2
CACHE_INSERTED = 1
CACHE_REUSED = 2

P8 = C.POINTER(C.c_uint8)
vp = C.c_void_p


class Config(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("device_ordinal", C.c_int32),
        ("queue_depth", C.c_uint32),
        ("hbm_budget_bytes", C.c_uint64),
        ("persistent_groups", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class MjpegFrame(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("component_count", C.c_uint32),
        ("h_sampling", C.c_uint8 * 4),
        ("v_sampling", C.c_uint8 * 4),
        ("quant_table", C.c_uint8 * 4),
        ("block_count", C.c_uint32),
        ("levels", vp),
        ("quant_tables", vp),
        ("planes", vp * 4),
        ("rgb", vp),
        ("rgb_range", C.c_uint32),
    ]


class Mpeg2Macroblock(C.Structure):
    _fields_ = [
        ("flags", C.c_uint8),
        ("cbp", C.c_uint8),
        ("qscale", C.c_uint8),
        ("reserved", C.c_uint8),
        ("mv", (C.c_int16 * 2) * 2),
        ("block_offset", C.c_uint32),
    ]


class Mpeg2Picture(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("picture_type", C.c_uint32),
        ("alternate_scan", C.c_uint32),
        ("intra_dc_precision", C.c_uint32),
        ("matrices", vp),
        ("macroblocks", vp),
        ("levels", vp),
        ("planes", vp * 3),
        ("forward", vp * 3),
        ("backward", vp * 3),
        ("forward_source", C.c_int32),
        ("backward_source", C.c_int32),
        ("field_motion", vp),
    ]


class Mpeg2FieldMotion(C.Structure):
    _fields_ = [("mv", (C.c_int16 * 2) * 2), ("field_select", (C.c_uint8 * 2) * 2)]


class DvFrame(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("system", C.c_uint32),
        ("levels", vp),
        ("block_info", vp),
        ("planes", vp * 3),
    ]


class H264Macroblock(C.Structure):
    _fields_ = [
        ("flags", C.c_uint8),
        ("qp", C.c_uint8),
        ("cbp", C.c_uint8),
        ("reserved", C.c_uint8),
        ("slice_id", C.c_uint16),
        ("offset_a", C.c_int8),
        ("offset_b", C.c_int8),
        ("coef_offset", C.c_uint32),
    ]


class H264Block(C.Structure):
    _fields_ = [
        ("mv", (C.c_int16 * 2) * 2),
        ("ref_idx", C.c_int8 * 2),
        ("total_coeff", C.c_uint8),
        ("reserved", C.c_uint8),
    ]


class H264Weight(C.Structure):
    _fields_ = [
        ("weight", C.c_int16),
        ("offset", C.c_int16),
        ("log2_denom", C.c_uint8),
        ("reserved", C.c_uint8 * 3),
    ]


class H264BiWeight(C.Structure):
    _fields_ = [
        ("w0", C.c_int16),
        ("w1", C.c_int16),
        ("o0", C.c_int16),
        ("o1", C.c_int16),
        ("log2_denom", C.c_uint8),
        ("reserved", C.c_uint8 * 3),
    ]


class H264Intra(C.Structure):
    _fields_ = [
        ("type", C.c_uint8),
        ("luma_mode", C.c_uint8),
        ("chroma_mode", C.c_uint8),
        ("reserved", C.c_uint8),
        ("modes", C.c_uint8 * 16),
    ]


class H264Reference(C.Structure):
    _fields_ = [
        ("planes", vp * 3),
        ("poc", C.c_int32),
        ("reserved", C.c_int32),
    ]


class H264Picture(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("width_mbs", C.c_uint32),
        ("height_mbs", C.c_uint32),
        ("weighted_mode", C.c_uint32),
        ("deblock", C.c_uint32),
        ("chroma_qp_offset", C.c_int32 * 2),
        ("macroblocks", vp),
        ("blocks", vp),
        ("levels", vp),
        ("weights", vp),
        ("bi_weights", vp),
        ("scaling_lists", vp),
        ("planes", vp * 3),
        ("ref_count", C.c_uint32 * 2),
        ("refs", (H264Reference * H264_MAX_REFS) * 2),
        ("intra", vp),
        ("transform_bypass", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class YuvToRgbJob(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("chroma_width", C.c_uint32),
        ("chroma_height", C.c_uint32),
        ("luma_stride", C.c_uint32),
        ("chroma_stride", C.c_uint32),
        ("range", C.c_uint32),
        ("matrix", C.c_uint32),
        ("planes", vp * 3),
        ("rgb", vp),
        ("rgba", vp),
    ]


class Layer(C.Structure):
    _fields_ = [
        ("rgba", vp),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("stride", C.c_uint32),
        ("x", C.c_uint32),
        ("y", C.c_uint32),
        ("op", C.c_uint32),
        ("opacity", C.c_uint32),
    ]


class CompositeJob(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("target", vp),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("stride", C.c_uint32),
        ("layer_count", C.c_uint32),
        ("layers", Layer * MAX_LAYERS),
    ]


class LinearLayer(C.Structure):
    _fields_ = [
        ("pixels", vp),
        ("format", C.c_uint32),
        ("stride", C.c_uint32),
        ("x", C.c_int32),
        ("y", C.c_int32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("clip_left", C.c_int32),
        ("clip_top", C.c_int32),
        ("clip_right", C.c_int32),
        ("clip_bottom", C.c_int32),
        ("opacity", C.c_uint32),
    ]


class LinearCompositeJob(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("target", vp),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("stride", C.c_uint32),
        ("layer_count", C.c_uint32),
        ("layers", LinearLayer * MAX_LAYERS),
    ]


class LinearConvertJob(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("direction", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("linear", vp),
        ("linear_stride", C.c_uint32),
        ("rgba_stride", C.c_uint32),
        ("rgba", vp),
    ]


class FrameKey(C.Structure):
    _fields_ = [
        ("namespace_id", C.c_uint32),
        ("variant", C.c_uint32),
        ("owner", C.c_uint64),
        ("revision", C.c_uint64),
        ("local_frame", C.c_int64),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


class FrameDescriptor(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("width", "height", "format", "color", "alpha_mode")]


class FrameCacheStats(C.Structure):
    _fields_ = [("struct_size", C.c_size_t)] + [(n, C.c_uint64) for n in (
        "resources", "retained_bytes", "byte_budget", "reuses", "insertions", "evictions", "generation")]


class YuvBlendJob(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("target", vp * 3),
        ("target_width", C.c_uint32),
        ("target_height", C.c_uint32),
        ("target_stride", C.c_uint32 * 3),
        ("overlay", vp * 3),
        ("rgba", vp),
        ("chroma_alpha", vp),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("x", C.c_uint32),
        ("y", C.c_uint32),
    ]


class OverlayPrepareJob(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("rgba", vp),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("y_limited", vp),
        ("y_full", vp),
        ("u_limited", vp),
        ("v_limited", vp),
        ("u_full", vp),
        ("v_full", vp),
        ("chroma_alpha", vp),
    ]


SCALE_FILTER_LANCZOS3 = 0
SCALE_FILTER_TRIANGLE = 1
SCALE_FIT = 0
SCALE_FILL = 1
SCALE_STRETCH = 2
SCALE_NATIVE = 3


class ScaleJob(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("src", vp * 3),
        ("src_stride", C.c_uint32 * 3),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("dst", vp * 3),
        ("dst_stride", C.c_uint32 * 3),
        ("canvas_width", C.c_uint32),
        ("canvas_height", C.c_uint32),
        ("mode", C.c_uint32),
        ("range", C.c_uint32),
        ("filter", C.c_uint32),
    ]


CODEC_MJPEG = 1
CODEC_MPEG2 = 2
CODEC_DV = 3
CODEC_H264 = 4


class ServerConfig(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("hbm_budget_bytes", C.c_uint64),
        ("slots", C.c_uint32),
        ("max_batch_frames", C.c_uint32),
        ("max_frames_per_request_per_batch", C.c_uint32),
        ("max_pending", C.c_uint32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("codec", C.c_uint32),
        ("stream_id", C.c_uint32),
        ("frame_count", C.c_uint32),
        ("frames", vp),
        ("convert", vp),
        ("composite", vp),
        ("hbm_bytes", C.c_uint64),
        ("user_tag", C.c_uint64),
    ]


class FrameCompletion(C.Structure):
    _fields_ = [
        ("request_id", C.c_uint64),
        ("user_tag", C.c_uint64),
        ("frame_index", C.c_uint32),
        ("status", C.c_int32),
        ("latency_ms", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("requests_enqueued", C.c_uint64),
        ("requests_completed", C.c_uint64),
        ("frames_completed", C.c_uint64),
        ("batches_launched", C.c_uint64),
        ("admitted_bytes", C.c_uint64),
        ("peak_admitted_bytes", C.c_uint64),
        ("pending", C.c_uint32),
        ("active", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("struct_size", C.c_size_t),
        ("mjpeg_group_blocks", C.c_uint32),
        ("mpeg2_group_blocks", C.c_uint32),
        ("dv_group_blocks", C.c_uint32),
        ("h264_deblock_band_rows", C.c_uint32),
        ("h264_deblock_groups_per_cu", C.c_uint32),
        ("h264_deblock_publish", C.c_uint32),
    ]


class TuningTrial(C.Structure):
    _fields_ = [("tuning", Tuning), ("kernel_ms", C.c_double)]


def new(struct_type, **fields):
    """Zero-initialised struct with struct_size set (if it has one) and fields assigned."""
    s = struct_type()
    if any(f[0] == "struct_size" for f in struct_type._fields_):
        s.struct_size = C.sizeof(struct_type)
    for k, v in fields.items():
        setattr(s, k, v)
    return s


# Exported entry points of libmmr_recon.so: (name, restype, argtypes).
ENTRY_POINTS = [
    ("mmr_recon_abi_version", C.c_uint32, []),
    ("mmr_recon_last_error_message", C.c_size_t, [C.c_char_p, C.c_size_t]),
    ("mmr_recon_context_create", C.c_int32, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    ("mmr_recon_context_destroy", None, [C.c_void_p]),
    ("mmr_recon_device_alloc", C.c_int32, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("mmr_recon_device_free", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mmr_recon_copy_to_device", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mmr_recon_copy_to_host", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mmr_recon_wait", C.c_int32, [C.c_void_p, C.c_uint64]),
    ("mmr_recon_poll", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32)]),
    ("mmr_recon_synchronize", C.c_int32, [C.c_void_p]),
    ("mmr_recon_last_kernel_ms", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]),
    ("mmr_recon_kernel_phases_ms", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.c_uint32,
                                               C.POINTER(C.c_uint32)]),
    ("mmr_recon_mjpeg_submit", C.c_int32,
     [C.c_void_p, C.POINTER(MjpegFrame), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_mpeg2_submit", C.c_int32,
     [C.c_void_p, C.POINTER(Mpeg2Picture), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_mpeg2_reconstruct_noalloc", C.c_int32, [C.c_void_p, C.POINTER(Mpeg2Picture), C.c_uint32, C.c_uint32]),
    ("mmr_recon_dv_submit", C.c_int32,
     [C.c_void_p, C.POINTER(DvFrame), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_h264_submit", C.c_int32,
     [C.c_void_p, C.POINTER(H264Picture), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_h264_implicit_weights", C.c_int32,
     [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("mmr_recon_yuv_to_rgb_submit", C.c_int32,
     [C.c_void_p, C.POINTER(YuvToRgbJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_composite_submit", C.c_int32,
     [C.c_void_p, C.POINTER(CompositeJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_yuv_composite_submit", C.c_int32,
     [C.c_void_p, C.POINTER(YuvToRgbJob), C.POINTER(CompositeJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_linear_composite_submit", C.c_int32,
     [C.c_void_p, C.POINTER(LinearCompositeJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_linear_convert_submit", C.c_int32,
     [C.c_void_p, C.POINTER(LinearConvertJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_frame_cache_create", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("mmr_recon_frame_cache_destroy", None, [C.c_void_p]),
    ("mmr_recon_frame_cache_begin_frame", C.c_int32, [C.c_void_p]),
    ("mmr_recon_frame_cache_retain", C.c_int32, [C.c_void_p, C.POINTER(FrameKey), C.POINTER(FrameDescriptor), C.c_uint64,
                                                 C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    ("mmr_recon_frame_cache_get", C.c_int32, [C.c_void_p, C.POINTER(FrameKey), C.POINTER(FrameDescriptor),
                                              C.POINTER(C.c_void_p)]),
    ("mmr_recon_frame_cache_remove", C.c_int32, [C.c_void_p, C.POINTER(FrameKey), C.POINTER(C.c_int32)]),
    ("mmr_recon_frame_cache_release_idle", C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("mmr_recon_frame_cache_clear", C.c_int32, [C.c_void_p]),
    ("mmr_recon_frame_cache_stats_get", C.c_int32, [C.c_void_p, C.POINTER(FrameCacheStats)]),
    ("mmr_recon_yuv_blend_submit", C.c_int32,
     [C.c_void_p, C.POINTER(YuvBlendJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_overlay_prepare_submit", C.c_int32,
     [C.c_void_p, C.POINTER(OverlayPrepareJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_scale_submit", C.c_int32,
     [C.c_void_p, C.POINTER(ScaleJob), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mmr_recon_server_create", C.c_int32, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    ("mmr_recon_server_enqueue", C.c_int32, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    ("mmr_recon_server_poll", C.c_int32,
     [C.c_void_p, C.POINTER(FrameCompletion), C.c_uint32, C.POINTER(C.c_uint32)]),
    ("mmr_recon_server_drain", C.c_int32, [C.c_void_p]),
    ("mmr_recon_server_stats_get", C.c_int32, [C.c_void_p, C.POINTER(ServerStats)]),
    ("mmr_recon_server_destroy", None, [C.c_void_p]),
    ("mmr_recon_prepare", C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("mmr_recon_prepared_launch", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    ("mmr_recon_prepared_destroy", None, [C.c_void_p, C.c_void_p]),
    ("mmr_recon_get_tuning", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mmr_recon_set_tuning", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("mmr_recon_autotune", C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
]
