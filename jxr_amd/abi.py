"""ctypes view of include/jxr_hip.h, jxr_hip_encode.h and jxr_hip_codestream.h.

The structures here mirror the C declarations field for field; a CPU test
compiles the header with gcc and checks every size and offset against them.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjxr_hip.so")

ABI_VERSION = 1

# status codes
OK = 0
ERR_INVALID_PLAN = 1
ERR_UNSUPPORTED_OUTPUT = 2
ERR_UNAVAILABLE = 3
ERR_DEVICE = 4
ERR_KERNEL_ARITHMETIC = 5
ERR_INVALID_ARGUMENT = 6
ERR_BUFFER_TOO_SMALL = 7
ERR_BUDGET = 8
ERR_SYNTAX = 9
ERR_UNSUPPORTED = 10
ERR_OVERFLOW = 11
TRIM_NONE = 0xFFFFFFFF

STAGE_FIRST_TRANSFORM = 1
STAGE_OVERLAP = 2
STAGE_PREDICTION = 3
STAGE_SECOND_TRANSFORM = 4
STAGE_SAMPLE_SCALE = 5
STAGE_OUTPUT = 16

BANDS_DC_ONLY, BANDS_NO_HIGH_PASS, BANDS_NO_FLEXBITS, BANDS_ALL = 0, 1, 2, 3
PRED_NONE, PRED_FROM_LEFT, PRED_FROM_TOP, PRED_FROM_TOP_LEFT = 0, 1, 2, 3

COLOR_LUMA, COLOR_YUV420, COLOR_YUV422, COLOR_YUV444, COLOR_RGB = 0, 1, 2, 3, 4
COLOR_CMYK, COLOR_CMYK_DIRECT, COLOR_YUVK, COLOR_RGBE, COLOR_NCOMPONENT = 5, 6, 7, 8, 9
ALPHA_NONE, ALPHA_INTEGRATED, ALPHA_SEPARATE = 0, 1, 2
SCALE_FULL, SCALE_QUARTER, SCALE_SIXTEENTH = 0, 1, 2

DEPTH_BIT1_WHITE, DEPTH_BIT1_BLACK, DEPTH_U8, DEPTH_U10, DEPTH_U16 = 0, 1, 2, 3, 4
DEPTH_I16, DEPTH_F16, DEPTH_I32, DEPTH_F32 = 5, 6, 7, 8
DEPTH_RGB555, DEPTH_RGB101010, DEPTH_RGB565 = 9, 10, 11

STORE_BITS, STORE_U8, STORE_U16, STORE_I16, STORE_I32 = 0, 1, 2, 3, 4
STORE_F16, STORE_F32, STORE_PACKED16, STORE_PACKED32 = 5, 6, 7, 8

LAYOUT_LUMA, LAYOUT_LUMA_ALPHA, LAYOUT_YUV, LAYOUT_YUVA = 0, 1, 2, 3
LAYOUT_RGB, LAYOUT_RGBA, LAYOUT_BGR, LAYOUT_BGRA = 4, 5, 6, 7
LAYOUT_CMYK, LAYOUT_CMYKA, LAYOUT_NCOMPONENT, LAYOUT_NCOMPONENT_ALPHA = 8, 9, 10, 11
LAYOUT_RGBX, LAYOUT_BGRX = 12, 13

BATCH_PER_IMAGE = 0
BATCH_PERSISTENT = 1

INPUT_U8, INPUT_U16, INPUT_F32 = 0, 1, 2

MAX_SURFACE_PLANES = 5

u32p = C.POINTER(C.c_uint32)


class CoefficientPlane(C.Structure):
    _fields_ = [
        ("coefficient_offset", C.c_uint64),
        ("coefficient_count", C.c_uint64),
        ("macroblock_offset", C.c_uint64),
        ("macroblock_count", C.c_uint64),
        ("block_columns", C.c_uint32),
        ("block_rows", C.c_uint32),
    ]


class Arena(C.Structure):
    _fields_ = [
        ("coefficients", C.POINTER(C.c_int32)),
        ("coefficient_count", C.c_uint64),
        ("macroblock_count", C.c_uint64),
        ("coefficient_offsets", u32p),
        ("quantizers", u32p),
        ("bands", C.POINTER(C.c_uint8)),
        ("hp_predictions", C.POINTER(C.c_uint8)),
        ("coded_x", u32p),
        ("coded_y", u32p),
        ("planes", C.POINTER(CoefficientPlane)),
        ("plane_count", C.c_uint32),
        ("overlap", C.c_uint32),
        ("hard_tiles", C.c_uint32),
        ("tile_column_count", C.c_uint32),
        ("tile_column_widths", u32p),
        ("tile_row_count", C.c_uint32),
        ("tile_row_heights", u32p),
        ("reserved0", C.c_uint32),
        ("identity", C.c_uint64),
    ]


class OutputFormat(C.Structure):
    _fields_ = [
        ("internal_color", C.c_uint32),
        ("output_color", C.c_uint32),
        ("bit_depth", C.c_uint32),
        ("shift_bits", C.c_uint32),
        ("mantissa_length", C.c_uint32),
        ("exponent_bias", C.c_int32),
        ("has_alpha_format", C.c_uint32),
        ("alpha_bit_depth", C.c_uint32),
        ("alpha_shift_bits", C.c_uint32),
        ("alpha_mantissa_length", C.c_uint32),
        ("alpha_exponent_bias", C.c_int32),
        ("alpha_scaled", C.c_uint32),
        ("storage", C.c_uint32),
        ("channel_layout", C.c_uint32),
        ("channel_count", C.c_uint32),
        ("scaled", C.c_uint32),
        ("red_blue_not_swapped", C.c_uint32),
        ("premultiply_alpha", C.c_uint32),
        ("chroma_centering_x", C.c_uint32),
        ("chroma_centering_y", C.c_uint32),
        ("crop_x", C.c_uint32),
        ("crop_y", C.c_uint32),
        ("crop_width", C.c_uint32),
        ("crop_height", C.c_uint32),
    ]


class SurfacePlane(C.Structure):
    _fields_ = [
        ("byte_offset", C.c_uint64),
        ("row_stride_bytes", C.c_uint64),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("channels", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class SurfaceLayout(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("plane_count", C.c_uint32),
        ("planar", C.c_uint32),
        ("planes", SurfacePlane * MAX_SURFACE_PLANES),
        ("byte_len", C.c_uint64),
    ]


class ImageDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("color_format", C.c_uint32),
        ("component_count", C.c_uint32),
        ("alpha_mode", C.c_uint32),
        ("scaled", C.c_uint32),
        ("scale", C.c_uint32),
        ("primary", C.POINTER(Arena)),
        ("separate_alpha", C.POINTER(Arena)),
        ("output", OutputFormat),
        ("surface", SurfaceLayout),
    ]


class Diagnostics(C.Structure):
    _fields_ = [
        ("upload_hits", C.c_uint64),
        ("upload_misses", C.c_uint64),
        ("upload_evictions", C.c_uint64),
        ("upload_resident_bytes", C.c_uint64),
        ("scratch_hits", C.c_uint64),
        ("scratch_misses", C.c_uint64),
        ("scratch_high_water_bytes", C.c_uint64),
        ("submissions", C.c_uint64),
        ("kernel_launches", C.c_uint64),
        ("stream_count", C.c_uint32),
        ("device_ordinal", C.c_int32),
        ("compute_units", C.c_uint32),
        ("reserved0", C.c_uint32),
    ]


class StageReport(C.Structure):
    _fields_ = [("stage", C.c_uint32), ("executor", C.c_uint32)]


class DecodeReport(C.Structure):
    _fields_ = [
        ("requested", C.c_uint32),
        ("selected", C.c_uint32),
        ("fallback", C.c_uint32),
        ("stage_count", C.c_uint32),
        ("stages", StageReport * 14),
    ]


# DecodeStage / StageExecutor / BackendKind codes (include/jxr_hip.h)
DECODE_STAGES = ["Parse", "EntropyDecode", "InverseScan", "CoefficientRemap", "DcLowPassPrediction",
                 "DequantizeAndFirstInverseTransform", "FirstOverlap", "HighPassPrediction",
                 "SecondInverseTransform", "SecondOverlap", "ChromaReconstruction", "ColorAndAlphaConversion",
                 "CropClipAndPack", "HostReadback"]
EXECUTORS = ["CpuScalar", "CpuSimd", "Metal", "Hip"]
BACKEND_AUTO, BACKEND_CPU, BACKEND_HIP = 0, 1, 2


class TileChoice(C.Structure):
    _fields_ = [
        ("bucket_key", C.c_uint64),
        ("tile_mbs_x", C.c_uint32),
        ("tile_mbs_y", C.c_uint32),
        ("images", C.c_uint32),
        ("candidates", C.c_uint32),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
        ("groups_per_cu", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("max_inflight", C.c_uint32),
        ("max_wave_images", C.c_uint32),
        ("max_waves", C.c_uint32),
        ("reserved0", C.c_uint32),
    ]


class ServerResult(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("tag", C.c_uint64),
        ("status", C.c_int32),
        ("stage", C.c_uint32),
        ("image", C.c_void_p),
        ("latency_ms", C.c_double),
        ("queue_ms", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("admitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("rejected", C.c_uint64),
        ("waves", C.c_uint64),
        ("max_wave_images", C.c_uint64),
        ("pixels_completed", C.c_uint64),
        ("inflight_high_water_bytes", C.c_uint64),
        ("hbm_high_water_bytes", C.c_uint64),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("queue_p50_ms", C.c_double),
        ("queue_p99_ms", C.c_double),
        ("mean_wave_images", C.c_double),
    ]


class ValuRate(C.Structure):
    _fields_ = [
        ("ms", C.c_double),
        ("wave_instructions", C.c_double),
        ("simds", C.c_uint32),
        ("clock_khz", C.c_uint32),
    ]


class EncodeParams(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("input_kind", C.c_uint32),
        ("input_channels", C.c_uint32),
        ("color_format", C.c_uint32),
        ("alpha_mode", C.c_uint32),
        ("overlap", C.c_uint32),
        ("hard_tiles", C.c_uint32),
        ("tile_column_count", C.c_uint32),
        ("tile_column_widths", u32p),
        ("tile_row_count", C.c_uint32),
        ("tile_row_heights", u32p),
        ("scaled", C.c_uint32),
        ("qp_dc", C.c_uint32),
        ("qp_lp", C.c_uint32),
        ("qp_hp", C.c_uint32),
        ("bands", C.c_uint32),
        ("shift_bits", C.c_uint32),
        ("mantissa_length", C.c_uint32),
        ("exponent_bias", C.c_int32),
        ("alpha_qp", C.c_uint32),
        ("seed", C.c_uint32),
        ("vary_quantizers", C.c_uint32),
        ("chroma_qp_set", C.c_uint32),
        ("qp_chroma_dc", C.c_uint32),
        ("qp_chroma_lp", C.c_uint32),
        ("qp_chroma_hp", C.c_uint32),
    ]


class CodestreamInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "width", "height", "margin_top", "margin_left", "margin_bottom", "margin_right",
        "macroblocks_x", "macroblocks_y", "tile_column_count", "tile_row_count", "overlap", "hard_tiles",
        "frequency_mode", "orientation", "output_color", "bit_depth", "color_format", "component_count",
        "bands", "scaled", "chroma_centering_x", "chroma_centering_y", "shift_bits", "mantissa_length")] + [
        ("exponent_bias", C.c_int32)] + [(n, C.c_uint32) for n in (
            "alpha_mode", "alpha_bands", "alpha_scaled", "alpha_shift_bits", "alpha_mantissa_length")] + [
        ("alpha_exponent_bias", C.c_int32)] + [(n, C.c_uint32) for n in (
            "alpha_bit_depth", "red_blue_not_swapped", "premultiplied_alpha", "short_header", "long_word",
            "trim_flexbits_present", "index_table_present", "profile_idc", "level_idc", "annex_a")] + [
        ("pixel_format_guid", C.c_uint8 * 16),
        ("codestream_offset", C.c_uint64),
        ("codestream_bytes", C.c_uint64),
        ("alpha_offset", C.c_uint64),
        ("alpha_bytes", C.c_uint64),
    ]


class CodestreamOptions(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "frequency_mode", "trim_flexbits", "annex_a", "short_header", "profile_idc", "level_idc")]


vp = C.c_void_p
pp = C.POINTER(C.c_void_p)
st = C.c_int32

# name -> (restype, argtypes); every symbol the headers declare
SIGNATURES = {
    "jxrhip_abi_version": (C.c_uint32, []),
    "jxrhip_last_error": (C.c_char_p, []),
    "jxrhip_is_available": (C.c_int32, []),
    "jxrhip_route_threshold": (C.c_uint64, []),
    "jxrhip_route": (st, [C.c_uint64, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "jxrhip_plan_create": (st, [C.POINTER(ImageDesc), pp]),
    "jxrhip_plan_destroy": (None, [vp]),
    "jxrhip_plan_output_bytes": (C.c_uint64, [vp]),
    "jxrhip_plan_reconstructed_coefficients": (C.c_uint64, [vp]),
    "jxrhip_plan_output_pixels": (C.c_uint64, [vp]),
    "jxrhip_plan_device_footprint": (C.c_uint64, [vp]),
    "jxrhip_plan_algorithmic_bytes": (C.c_uint64, [vp]),
    "jxrhip_session_create": (st, [C.c_int32, pp]),
    "jxrhip_session_destroy": (None, [vp]),
    "jxrhip_session_diagnostics": (st, [vp, C.POINTER(Diagnostics)]),
    "jxrhip_session_set_upload_budget": (st, [vp, C.c_uint64]),
    "jxrhip_session_upload": (st, [vp, vp]),
    "jxrhip_session_clear_cache": (st, [vp]),
    "jxrhip_submit": (st, [vp, vp, pp]),
    "jxrhip_submit_into": (st, [vp, vp, vp, C.c_uint64, pp]),
    "jxrhip_submission_is_complete": (C.c_int32, [vp]),
    "jxrhip_submission_wait": (st, [vp, pp, u32p]),
    "jxrhip_submission_destroy": (None, [vp]),
    "jxrhip_submission_enqueue_wait": (st, [vp, vp]),
    "jxrhip_submission_output_layout": (st, [vp, C.POINTER(SurfaceLayout)]),
    "jxrhip_submission_pending_device_ptr": (st, [vp, pp, C.POINTER(C.c_uint64)]),
    "jxrhip_plan_decode_report": (st, [vp, C.c_uint32, C.c_int32, C.POINTER(DecodeReport)]),
    "jxrhip_batch_enqueue_wait": (st, [vp, vp]),
    "jxrhip_allocate_destination": (st, [vp, C.c_uint64, pp]),
    "jxrhip_free_destination": (st, [vp, vp]),
    "jxrhip_fill_destination": (st, [vp, vp, C.c_uint8, C.c_uint64]),
    "jxrhip_copy_to_host": (st, [vp, vp, vp, C.c_uint64]),
    "jxrhip_session_synchronize": (st, [vp]),
    "jxrhip_copy_device": (st, [vp, vp, vp, C.c_uint64]),
    "jxrhip_image_device_ptr": (vp, [vp]),
    "jxrhip_image_bytes": (C.c_uint64, [vp]),
    "jxrhip_image_destroy": (None, [vp]),
    "jxrhip_readback": (st, [vp, vp, vp, C.c_uint64]),
    "jxrhip_decode_to_host": (st, [vp, vp, vp, C.c_uint64, u32p]),
    "jxrhip_submit_batch": (st, [vp, C.POINTER(vp), C.c_uint32, C.POINTER(vp), C.c_uint32, pp]),
    "jxrhip_decode_batch_static": (st, [vp, C.POINTER(vp), C.c_uint32, C.POINTER(vp), u32p]),
    "jxrhip_batch_is_complete": (C.c_int32, [vp]),
    "jxrhip_batch_wait": (st, [vp, u32p]),
    "jxrhip_last_batch_timing": (st, [vp, C.POINTER(C.c_double), u32p]),
    "jxrhip_session_set_tile": (st, [vp, C.c_uint32, C.c_uint32]),
    "jxrhip_session_set_occupancy": (st, [vp, C.c_uint32]),
    "jxrhip_plan_bucket_key": (C.c_uint64, [vp]),
    "jxrhip_session_autotune": (st, [vp, C.POINTER(vp), C.c_uint32, C.POINTER(vp), C.c_uint32,
                                     C.POINTER(TileChoice), C.c_uint32, u32p]),
    "jxrhip_session_set_bucket_tile": (st, [vp, C.c_uint64, C.c_uint32, C.c_uint32]),
    "jxrhip_session_bucket_tile": (st, [vp, C.c_uint64, u32p, u32p]),
    "jxrhip_batch_graph_create": (st, [vp, C.POINTER(vp), C.c_uint32, C.POINTER(vp), pp]),
    "jxrhip_batch_graph_launch": (st, [vp, pp]),
    "jxrhip_batch_graph_destroy": (None, [vp]),
    "jxrhip_server_create": (st, [vp, C.POINTER(ServerConfig), pp]),
    "jxrhip_server_clock_ns": (C.c_uint64, []),
    "jxrhip_server_submit": (st, [vp, vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]),
    "jxrhip_server_poll": (C.c_uint32, [vp, C.POINTER(ServerResult), C.c_uint32, C.c_uint32]),
    "jxrhip_server_drain": (st, [vp]),
    "jxrhip_server_get_stats": (st, [vp, C.POINTER(ServerStats)]),
    "jxrhip_server_destroy": (None, [vp]),
    # jxr_hip_encode.h
    "jxrhip_encode": (st, [vp, C.c_uint64, C.POINTER(EncodeParams), pp]),
    "jxrhip_encoded_destroy": (None, [vp]),
    "jxrhip_encoded_arena": (C.POINTER(Arena), [vp, C.c_uint32]),
    "jxrhip_quant_step": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "jxrhip_calibrate_valu": (st, [vp, C.c_uint32, C.POINTER(ValuRate)]),
    "jxrhip_quality_parameters": (st, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p,
                                      C.POINTER(C.c_uint32)]),
    "jxrhip_float32_sample": (st, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_int32)]),
    "jxrhip_quality_parameters_scaled": (st, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_uint32, u32p, u32p, u32p]),
    "jxrhip_sample_scale": (st, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, u32p, u32p,
                                 u32p]),
    # jxr_hip_codestream.h
    "jxrhip_codestream_parse": (st, [vp, C.c_uint64, pp]),
    "jxrhip_codestream_destroy": (None, [vp]),
    "jxrhip_codestream_get_info": (st, [vp, C.POINTER(CodestreamInfo)]),
    "jxrhip_codestream_tile_grid": (st, [vp, u32p, u32p]),
    "jxrhip_codestream_decode": (st, [vp, C.c_uint32, vp, C.c_uint32, pp]),
    "jxrhip_decoded_destroy": (None, [vp]),
    "jxrhip_codestream_coefficient_count": (st, [vp, C.c_uint32, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "jxrhip_codestream_decode_into": (st, [vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint64, vp, C.c_uint64, pp]),
    "jxrhip_host_alloc": (st, [vp, C.c_uint64, pp]),
    "jxrhip_host_free": (st, [vp, vp]),
    "jxrhip_decoded_arena": (C.POINTER(Arena), [vp, C.c_uint32]),
    "jxrhip_encoded_codestream": (st, [vp, C.POINTER(CodestreamOptions), C.POINTER(C.POINTER(C.c_uint8)),
                                       C.POINTER(C.c_uint64)]),
    "jxrhip_free_bytes": (None, [vp]),
}

_lib = None


class JxrHipError(RuntimeError):
    def __init__(self, code, message, stage=0):
        super().__init__(f"jxr_hip error {code}: {message}" + (f" (stage {stage})" if stage else ""))
        self.code = code
        self.stage = stage
        self.message = message


def load():
    """Loads libjxr_hip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.jxrhip_abi_version() != ABI_VERSION:
        raise ImportError("libjxr_hip.so ABI version differs from the Python bindings")
    _lib = lib
    return lib


def check(code, what=""):
    if code != OK:
        msg = load().jxrhip_last_error().decode(errors="replace")
        raise JxrHipError(code, f"{what}: {msg}" if what else msg)
