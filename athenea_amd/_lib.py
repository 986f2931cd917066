"""ctypes binding of libathx.so, the C ABI declared in include/athx.h.

The library is the product: nothing here computes a frame. Loading fails
loudly when the library has not been built (run ``python -m athenea_amd.build``
or ``__graft_entry__.build()``); there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

# ATHX_LIBRARY names another build of the same library (to compare builds).
LIB_PATH = Path(os.environ.get("ATHX_LIBRARY") or Path(__file__).resolve().parent / "libathx.so")

# athx_status: 0 = OK, else athenea's ErrorCode + 1 (ErrorCode numbers Unknown 0 ... InternalError 10,
# modules/core/include/athenea/core/Error.h:14-26); athx_error_code() gives the ErrorCode value.
STATUS_NAMES = {
    0: "Ok",
    1: "Unknown",
    2: "NotFound",
    3: "InvalidArgument",
    4: "Unsupported",
    5: "IoFailure",
    6: "DeviceFailure",
    7: "ShaderFailure",
    8: "PluginFailure",
    9: "OutOfMemory",
    10: "Cancelled",
    11: "InternalError",
}

CULL_REASONS = 7

c_float_p = ctypes.POINTER(ctypes.c_float)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)


class View(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("focal_x", ctypes.c_float),
        ("focal_y", ctypes.c_float),
        ("centre_x", ctypes.c_float),
        ("centre_y", ctypes.c_float),
        ("near_z", ctypes.c_float),
        ("far_z", ctypes.c_float),
        ("orthographic", ctypes.c_uint32),
        ("object_to_view", ctypes.c_float * 12),
        ("eye", ctypes.c_float * 3),
        ("background", ctypes.c_float * 4),
    ]


class Sky(ctypes.Structure):
    _fields_ = [
        ("texels", ctypes.c_void_p),
        ("sh", ctypes.c_void_p),
        ("sun", ctypes.c_void_p),
        ("base_side", ctypes.c_uint32),
    ]


class Relight(ctypes.Structure):
    _fields_ = [
        ("normals", ctypes.c_void_p),
        ("albedo", ctypes.c_void_p),
        ("metallic", ctypes.c_void_p),
        ("roughness", ctypes.c_void_p),
        ("normals_packed", ctypes.c_void_p),
        ("pbr", ctypes.c_void_p),
        ("object_to_world", ctypes.c_float * 12),
        ("sky", Sky),
    ]


class PackedSplats(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p),
        ("shape", ctypes.c_void_p),
        ("sh", ctypes.c_void_p),
        ("rest_per_colour", ctypes.c_uint32),
    ]


class Gaussians(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint32),
        ("sh_coeffs", ctypes.c_uint32),
        ("means", ctypes.c_void_p),
        ("log_scales", ctypes.c_void_p),
        ("quats", ctypes.c_void_p),
        ("opacity_logits", ctypes.c_void_p),
        ("sh", ctypes.c_void_p),
        ("sh_half", ctypes.c_void_p),
        ("relight", ctypes.POINTER(Relight)),
        ("packed", ctypes.POINTER(PackedSplats)),
        ("crypto_ids", ctypes.c_void_p),
    ]


class Instance(ctypes.Structure):
    _fields_ = [
        ("gaussians", ctypes.POINTER(Gaussians)),
        ("object_to_view", ctypes.c_void_p),
        ("eye", ctypes.c_void_p),
    ]


class Settings(ctypes.Structure):
    _fields_ = [
        ("tile", ctypes.c_uint32),
        ("batch", ctypes.c_uint32),
        ("antialias", ctypes.c_uint32),
        ("linear", ctypes.c_uint32),
        ("max_sh_degree", ctypes.c_uint32),
        ("depth_mode", ctypes.c_uint32),
        ("depth_threshold", ctypes.c_float),
        ("time_stages", ctypes.c_uint32),
        ("count_splats", ctypes.c_uint32),
        ("counters_tag", ctypes.c_uint64),
    ]


TRACE_SLOTS = 4
TRACE_WORDS = TRACE_SLOTS * 2 + TRACE_SLOTS // 2 + 1
TRACE_NONE = 0xFFFFFFFF
TRACE_EMPTY = 0xFFFFFFFE


class Targets(ctypes.Structure):
    _fields_ = [
        ("colour", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("final_t", ctypes.c_void_p),
        ("last", ctypes.c_void_p),
        ("crypto", ctypes.c_void_p),
        ("trace", ctypes.c_void_p),
        ("under_colour", ctypes.c_void_p),
        ("under_depth", ctypes.c_void_p),
        ("under_crypto", ctypes.c_void_p),
    ]


class FrameStats(ctypes.Structure):
    _fields_ = [
        ("splats", ctypes.c_uint32),
        ("visible", ctypes.c_uint32),
        ("pairs", ctypes.c_uint64),
        ("project_ms", ctypes.c_double),
        ("depth_sort_ms", ctypes.c_double),
        ("counts_ms", ctypes.c_double),
        ("emit_ms", ctypes.c_double),
        ("tile_sort_ms", ctypes.c_double),
        ("ranges_ms", ctypes.c_double),
        ("blend_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("frame_serial", ctypes.c_uint64),
    ]


MAX_COUNTED_CLOUDS = 64


class Counters(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_uint64),
        ("slots", ctypes.c_uint32),
        ("visible", ctypes.c_uint32),
        ("pairs", ctypes.c_uint64),
        ("max_tiles", ctypes.c_uint32),
        ("culled", ctypes.c_uint32 * CULL_REASONS),
        ("clouds", ctypes.c_uint32),
        ("cloud_visible", ctypes.c_uint32 * MAX_COUNTED_CLOUDS),
        ("cloud_pairs", ctypes.c_uint32 * MAX_COUNTED_CLOUDS),
    ]


class Grads(ctypes.Structure):
    _fields_ = [
        ("means", ctypes.c_void_p),
        ("log_scales", ctypes.c_void_p),
        ("quats", ctypes.c_void_p),
        ("opacity_logits", ctypes.c_void_p),
        ("sh", ctypes.c_void_p),
        ("overwrite", ctypes.c_uint32),
    ]


class FrameInfo(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("tiles", ctypes.c_uint32),
        ("key_bits", ctypes.c_uint32),
        ("pairs", ctypes.c_uint64),
        ("strips", ctypes.c_uint32),
        ("entries", ctypes.c_uint64),
    ]


# athx_frame_part
FRAME_PARTS = {"records": 0, "touched": 1, "depth_keys": 2, "offsets": 3, "pair_keys": 4, "pair_values": 5,
               "ranges": 6}


# Every symbol include/athx.h declares, with its signature.
PLY_MAX_PROPERTIES = 256
# athx_ply_type: the property types a vertex may hold.
PLY_TYPES = {"f32": 0, "f64": 1, "u8": 2, "i8": 3, "u16": 4, "i16": 5, "u32": 6, "i32": 7}


class PlyLayout(ctypes.Structure):
    _fields_ = [
        ("data_offset", ctypes.c_uint64),
        ("count", ctypes.c_uint64),
        ("stride", ctypes.c_uint32),
        ("properties", ctypes.c_uint32),
        ("offset", ctypes.c_uint32 * PLY_MAX_PROPERTIES),
        ("type", ctypes.c_uint32 * PLY_MAX_PROPERTIES),
        ("x", ctypes.c_uint32),
        ("y", ctypes.c_uint32),
        ("z", ctypes.c_uint32),
        ("opacity", ctypes.c_uint32),
        ("scale", ctypes.c_uint32 * 3),
        ("rot", ctypes.c_uint32 * 4),
        ("dc", ctypes.c_uint32 * 3),
        ("rest_base", ctypes.c_uint32),
        ("rest_per_colour", ctypes.c_uint32),
        ("opacity_mode", ctypes.c_uint32),
        ("scale_mode", ctypes.c_uint32),
        ("colour_mode", ctypes.c_uint32),
        ("rotation_mode", ctypes.c_uint32),
    ]


class IngestOut(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p),
        ("shape", ctypes.c_void_p),
        ("sh", ctypes.c_void_p),
        ("origin", ctypes.c_void_p),
    ]


class ImageDifference(ctypes.Structure):
    _fields_ = [("pixels", ctypes.c_uint64), ("p99", ctypes.c_uint32), ("max", ctypes.c_uint32),
                ("over2", ctypes.c_uint64)]


class HdrDifference(ctypes.Structure):
    _fields_ = [("pixels", ctypes.c_uint64), ("rel_mse", ctypes.c_double), ("p99_relative", ctypes.c_double),
                ("max_relative", ctypes.c_double)]


SIGNATURES = {
    "athx_version": (ctypes.c_char_p, []),
    "athx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "athx_destroy": (None, [ctypes.c_void_p]),
    "athx_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "athx_release_scratch": (ctypes.c_int, [ctypes.c_void_p]),
    "athx_scratch_bytes": (ctypes.c_uint64, [ctypes.c_void_p]),
    "athx_render": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            ctypes.POINTER(View),
            ctypes.c_uint32,
            ctypes.POINTER(Gaussians),
            ctypes.POINTER(Settings),
            ctypes.POINTER(Targets),
            ctypes.c_void_p,
            ctypes.POINTER(FrameStats),
        ],
    ),
    "athx_render_backward": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            ctypes.POINTER(View),
            ctypes.c_uint32,
            ctypes.POINTER(Gaussians),
            ctypes.POINTER(Settings),
            ctypes.POINTER(Targets),
            ctypes.c_uint64,
            ctypes.POINTER(ctypes.c_void_p),
            ctypes.POINTER(Grads),
            ctypes.c_void_p,
        ],
    ),
    "athx_frame_serial": (ctypes.c_uint64, [ctypes.c_void_p]),
    "athx_error_code": (ctypes.c_int, [ctypes.c_int]),
    "athx_render_async": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            ctypes.POINTER(View),
            ctypes.c_uint32,
            ctypes.POINTER(Gaussians),
            ctypes.POINTER(Settings),
            ctypes.POINTER(Targets),
            ctypes.c_uint64,
            ctypes.c_void_p,
        ],
    ),
    "athx_frame_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "athx_graph_create": (
        ctypes.c_int,
        [
            ctypes.c_int,
            ctypes.POINTER(View),
            ctypes.c_uint32,
            ctypes.POINTER(Gaussians),
            ctypes.POINTER(Settings),
            ctypes.POINTER(Targets),
            ctypes.c_uint64,
            ctypes.POINTER(ctypes.c_void_p),
            ctypes.POINTER(Grads),
            ctypes.POINTER(ctypes.c_void_p),
            ctypes.c_char_p,
            ctypes.c_uint32,
        ],
    ),
    "athx_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "athx_graph_launch_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "athx_graph_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "athx_graph_destroy": (None, [ctypes.c_void_p]),
    "athx_latest_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Counters)]),
    "athx_frame_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FrameInfo)]),
    "athx_copy_frame": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                       ctypes.c_void_p]),
    "athx_exclusive_scan": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p],
    ),
    "athx_sort_pairs": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p],
    ),
    "athx_ply_parse": (
        ctypes.c_int,
        [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(PlyLayout), ctypes.c_char_p, ctypes.c_uint64],
    ),
    "athx_contribution": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(View), ctypes.c_uint32, ctypes.POINTER(Gaussians), ctypes.POINTER(Settings),
         ctypes.POINTER(Targets), ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p],
    ),
    "athx_sky_texels": (ctypes.c_uint32, [ctypes.c_uint32]),
    "athx_prepare_sky": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
         ctypes.c_void_p],
    ),
    "athx_sky_sun": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_void_p]),
    "athx_sky_residual": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "athx_exclusive_scan_noalloc": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p],
    ),
    "athx_splat_layout": (
        ctypes.c_int,
        [ctypes.c_uint64, ctypes.POINTER(PlyLayout), ctypes.c_char_p, ctypes.c_uint64],
    ),
    "athx_render_instances": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(View), ctypes.c_uint32, ctypes.POINTER(Instance), ctypes.c_uint32,
         ctypes.POINTER(Settings), ctypes.POINTER(Targets), ctypes.c_void_p, ctypes.POINTER(FrameStats),
         ctypes.POINTER(ctypes.c_uint32)],
    ),
    "athx_reference_render_instances": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(View), ctypes.POINTER(Instance), ctypes.c_uint32, ctypes.POINTER(Settings),
         ctypes.POINTER(Targets), ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)],
    ),
    "athx_reference_render": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(View), ctypes.POINTER(Gaussians), ctypes.POINTER(Settings),
         ctypes.POINTER(Targets), ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)],
    ),
    "athx_compare_images": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
         ctypes.POINTER(ImageDifference)],
    ),
    "athx_compare_hdr": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
         ctypes.POINTER(HdrDifference)],
    ),
    "athx_ingest_ply": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(PlyLayout), ctypes.c_uint32, ctypes.c_uint32,
         ctypes.POINTER(IngestOut), ctypes.POINTER(ctypes.c_uint32)],
    ),
}

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


class AthxError(RuntimeError):
    """A failed call of the C ABI: ``code`` is the athenea ErrorCode name,
    ``error_code`` its ErrorCode value (status - 1)."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.code = STATUS_NAMES.get(status, f"status {status}")
        self.error_code = status - 1 if 1 <= status <= 11 else -1
        super().__init__(f"{self.code}: {message}")


def load() -> ctypes.CDLL:
    """Loads libathx.so once; raises if it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is not built: the rasterizer has no CPU path. "
                    "Build it with `python -m athenea_amd.build`."
                )
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = ""
        if ctx is not None:
            raw = load().athx_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise AthxError(status, msg)
