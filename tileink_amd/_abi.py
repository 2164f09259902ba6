"""ctypes view of include/tileink_hip.h and the loader for libtileink_hip.so.

The shared library is the product: it carries the gfx950 code object and the
host recorder.  Nothing here falls back to another implementation; a missing or
unloadable library raises immediately.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtileink_hip.so")

TK_OK, TK_ERR_INVALID, TK_ERR_HIP, TK_ERR_OOM, TK_ERR_UNAVAILABLE, TK_ERR_OVERFLOW = range(6)
STATUS_NAMES = {0: "ok", 1: "invalid", 2: "hip", 3: "oom", 4: "unavailable", 5: "overflow"}
INVALID = 0xFFFFFFFF
# tk_coarse_no_alloc's fixed scratch (include/tileink_hip.h)
TK_NO_ALLOC_SCRATCH_BYTES = 1 << 20
TK_NO_ALLOC_MAX_TABLE = 536_739_840

PATH_MOVE, PATH_LINE, PATH_QUAD, PATH_CUBIC, PATH_CLOSE = range(5)
GRADIENT_TRANSFORM, GRADIENT_UNPREMULTIPLIED = 1, 2
FILL_NONZERO, FILL_EVEN_ODD = 0, 1
DRAW_BRUSH, DRAW_CLIP = 0, 1
BRUSH_SOLID, BRUSH_LINEAR, BRUSH_RADIAL, BRUSH_SWEEP, BRUSH_FOUR_CORNER, BRUSH_PATTERN, BRUSH_IMAGE = 1, 2, 3, 4, 5, 6, 7
PATTERN_NEAREST, PATTERN_BILINEAR = 0, 1

# Frame buffer identifiers (tk_frame_buffer).
BUF_BACKDROPS, BUF_SEGMENT_COUNTS, BUF_SEGMENT_RANGES, BUF_SEGMENT_CURSORS, BUF_SEGMENT_BUMPS, BUF_SEGMENTS = range(6)
TK_TILE_KIND_INTERPRETER, TK_TILE_KIND_EMPTY, TK_TILE_KIND_COLOR, TK_TILE_KIND_SDF, TK_TILE_KIND_MIXED = range(5)
BUF_TILE_PTCL_COUNTS, BUF_TILE_PTCL_RANGES, BUF_TILE_KINDS, BUF_BATCH_PTCL_BASES, BUF_PTCL, BUF_RAW_BACKDROPS = range(6, 12)
(BUF_LINES, BUF_PATHS, BUF_DRAWS, BUF_TILE_DRAWS, BUF_TILE_DRAW_INDICES, BUF_SCAN_CHUNKS, BUF_SCAN_CHUNK_RANGES,
 BUF_CUMSUM_CHUNK_OFFSETS, BUF_CUMSUM_CHUNK_LENS, BUF_CUMSUM_ROW_STARTS, BUF_CUMSUM_ROW_ENDS) = range(12, 23)


class TileinkError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"tileink {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


# --- record structs (byte-exact with the header) ------------------------------

class Affine(C.Structure):
    _fields_ = [(n, C.c_float) for n in "abcdef"]


class Line(C.Structure):
    _fields_ = [("path_id", C.c_uint32), ("pad", C.c_float), ("p0", C.c_float * 2), ("p1", C.c_float * 2)]


class PathRecord(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("path_id", "line_count", "line_start", "flags", "data_offset", "data_len",
                                          "tile_x0", "tile_y0", "tile_x1", "tile_y1", "segment_start",
                                          "segment_capacity", "segment_count")] + [("transform", Affine)]


class PixelBounds(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("x0", "y0", "x1", "y1")]


class DrawRecord(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("path_id", "glyph_run_id", "sdf_offset", "sdf_len", "sdf_shadow_offset",
                                          "sdf_shadow_len", "brush_offset", "brush_len", "tag", "fill_rule")] + [
        ("pixel_bounds", PixelBounds), ("local_pixel_bounds", PixelBounds), ("solid_rect", C.c_uint32),
        ("transform", Affine), ("inverse_transform", Affine)]


class ScanConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("clear_len", "backdrop_len", "path_count", "scan_chunk_count", "line_count",
                                          "segment_capacity", "incremental", "line_base", "path_base", "chunk_base",
                                          "backdrop_base")]


class CumsumConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("row_count", "chunk_count", "pad0", "pad1")]


class Scene(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("tiles_width", C.c_uint32), ("tiles_height", C.c_uint32),
        ("clear_color", C.c_uint32),
        ("lines", C.c_void_p), ("line_count", C.c_uint32),
        ("paths", C.c_void_p), ("path_count", C.c_uint32),
        ("draws", C.c_void_p), ("draw_count", C.c_uint32),
        ("paint", C.c_void_p), ("paint_words", C.c_uint32),
        ("paint_sdf_shadow_base", C.c_uint32), ("paint_brush_base", C.c_uint32),
        ("draw_batch_ids", C.c_void_p),
        ("layers", C.c_void_p), ("layer_count", C.c_uint32),
        ("batches", C.c_void_p), ("batch_count", C.c_uint32),
        ("tile_draws", C.c_void_p),
        ("tile_draw_indices", C.c_void_p), ("tile_draw_index_count", C.c_uint32),
        ("scan_chunks", C.c_void_p), ("scan_chunk_count", C.c_uint32),
        ("scan_chunk_ranges", C.c_void_p),
        ("cumsum_chunk_offsets", C.c_void_p), ("cumsum_chunk_lens", C.c_void_p), ("cumsum_chunk_count", C.c_uint32),
        ("cumsum_row_starts", C.c_void_p), ("cumsum_row_ends", C.c_void_p), ("cumsum_row_count", C.c_uint32),
        ("backdrop_len", C.c_uint32), ("segment_capacity", C.c_uint32), ("max_clip_depth", C.c_uint32),
        ("ptcl_capacity", C.c_uint64),
        ("atlas", C.c_void_p), ("atlas_width", C.c_uint32), ("atlas_height", C.c_uint32), ("atlas_pages", C.c_uint32),
        ("atlas_reserved", C.c_uint32),
    ]


class GradientStop(C.Structure):
    _fields_ = [("offset", C.c_float), ("rgba", C.c_float * 4)]


class Brush(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("rgba", C.c_float * 4), ("start", C.c_float * 2), ("end", C.c_float * 2),
                ("extend", C.c_uint32), ("stops", C.POINTER(GradientStop)), ("stop_count", C.c_uint32),
                ("start_radius", C.c_float), ("end_radius", C.c_float), ("start_angle", C.c_float),
                ("end_angle", C.c_float), ("bounds", C.c_float * 4), ("corners", (C.c_float * 4) * 4),
                ("image", C.c_uint32), ("pattern_transform", C.c_float * 6), ("sampling", C.c_uint32),
                ("opacity", C.c_uint32), ("gradient_transform", C.c_float * 6), ("gradient_flags", C.c_uint32)]


class Limits(C.Structure):
    _fields_ = [("image_dimension", C.c_uint32), ("max_tiles", C.c_uint32), ("device_memory_bytes", C.c_uint64),
                ("compute_units", C.c_uint32)]


class FrameStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("paths", "lines", "draws", "batches", "tiles", "backdrop_len",
                                          "segment_capacity", "tile_draw_refs")] + [
        ("ptcl_capacity", C.c_uint64), ("ptcl_records", C.c_uint64), ("segments_emitted", C.c_uint32),
        ("overflow", C.c_uint32), ("device_bytes", C.c_uint64),
        ("scan_chunk_count", C.c_uint32), ("cumsum_chunk_count", C.c_uint32), ("cumsum_row_count", C.c_uint32),
        ("pad", C.c_uint32), ("ptcl_bound", C.c_uint64)]


class StageTimes(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("scan_clear", "scan_count", "scan_prefix", "scan_emit", "cumsum", "coarse_count", "coarse_prefix", "coarse_emit", "fine", "total")]


class StrokeStyle(C.Structure):
    _fields_ = [("width", C.c_double), ("miter_limit", C.c_double), ("join", C.c_uint32), ("start_cap", C.c_uint32),
                ("end_cap", C.c_uint32), ("reserved", C.c_uint32)]


TUNING_FIELDS = ("scan_xcd_run", "coarse_xcd_run", "fine_xcd_run", "persistent", "scan_block", "coarse_block",
                 "coarse_scan_items", "diff_keep_coverage")
TK_PERSISTENT_COARSE, TK_PERSISTENT_FINE = 1, 2


class Tuning(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in TUNING_FIELDS]


class CoarseConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("tile_count", "tiles_width", "tiles_height", "batch_count", "draw_count",
                                          "path_count", "paint_words", "paint_brush_base", "layer_count",
                                          "tile_draw_index_count", "pad0", "pad1")] + [("ptcl_capacity", C.c_uint64)]


class CoarseInputs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("draws", "paths", "backdrops", "segment_ranges", "layers", "batches",
                                          "paint", "draw_batch_ids", "tile_draws", "tile_draw_indices")]


class CoarseOutputs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("tile_ptcl_counts", "tile_ptcl_ranges", "tile_kinds", "batch_ptcl_bases",
                                          "ptcl", "status")]


class FineConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("width", "height", "clear_color", "tile_count", "tiles_width",
                                          "tiles_height", "batch_count", "clip_spill_depth", "paint_words",
                                          "paint_sdf_shadow_base", "paint_brush_base", "draw_count",
                                          "ptcl_capacity", "segment_capacity")]


class FineInputs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("draws", "paint", "segments", "tile_ptcl_ranges", "tile_kinds",
                                          "batch_ptcl_bases", "ptcl", "status", "atlas")] + \
               [(n, C.c_uint32) for n in ("atlas_width", "atlas_height", "atlas_pages", "atlas_reserved")]


# numpy dtypes for reading record arrays back
LINE_DTYPE = np.dtype([("path_id", "<u4"), ("pad", "<f4"), ("p0", "<f4", 2), ("p1", "<f4", 2)])
SEGMENT_DTYPE = np.dtype([("p0x", "<f4"), ("p0y", "<f4"), ("p1x", "<f4"), ("p1y", "<f4"), ("y_edge", "<f4")])
PTCL_DTYPE = np.dtype([("tag", "<u4"), ("backdrop", "<i4"), ("fill_rule", "<u4"), ("segment_start", "<u4"),
                       ("segment_end", "<u4"), ("color", "<u4")])
assert LINE_DTYPE.itemsize == C.sizeof(Line) == 24
assert C.sizeof(PathRecord) == 76 and C.sizeof(DrawRecord) == 124 and SEGMENT_DTYPE.itemsize == 20
assert PTCL_DTYPE.itemsize == 24 and C.sizeof(ScanConfig) == 44

# Every exported entry point, with (restype, argtypes).  The list is the C ABI.
_P = C.c_void_p
_U32 = C.c_uint32
_SZ = C.c_size_t
SIGNATURES = {
    "tk_last_error": (C.c_char_p, []),
    "tk_version": (_U32, []),
    "tk_canvas_create": (C.c_int, [_U32, _U32, C.c_double, C.POINTER(_P)]),
    "tk_canvas_destroy": (None, [_P]),
    "tk_canvas_push_path": (C.c_int, [_P, _P, _SZ, _P, _SZ, _P, _P, _U32, C.c_double, C.POINTER(_U32)]),
    "tk_canvas_push_text_outline": (C.c_int, [_P, _P, _SZ, _P, _SZ, _P, _P, C.c_double, C.POINTER(_U32)]),
    "tk_canvas_push_stroke": (C.c_int, [_P, _P, _SZ, _P, _SZ, C.c_double, _P, _P, _U32, C.c_double, C.POINTER(_U32)]),
    "tk_canvas_push_rect": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double, _P, _P, C.POINTER(_U32)]),
    "tk_canvas_add_image": (C.c_int, [_P, _U32, _U32, _P, C.POINTER(_U32)]),
    "tk_canvas_push_image": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double, _U32, _U32, _U32,
                                       C.POINTER(_U32)]),
    "tk_canvas_push_clip_sdf_rect_layer": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double, _P]),
    "tk_canvas_push_clip_layer": (C.c_int, [_P, _P, _SZ, _P, _SZ, _P, _U32, C.c_double]),
    "tk_canvas_pop_layer": (C.c_int, [_P]),
    "tk_canvas_encode": (C.c_int, [_P, _U32, C.POINTER(Scene)]),
    "tk_canvas_line_count": (_U32, [_P]),
    "tk_line_scanned_tile_count": (_U32, [C.POINTER(Line), _U32, _U32, _U32, _U32]),
    "tk_context_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "tk_context_destroy": (None, [_P]),
    "tk_context_stream": (_P, [_P]),
    "tk_context_synchronize": (C.c_int, [_P]),
    "tk_context_limits": (C.c_int, [_P, C.POINTER(Limits)]),
    "tk_device_alloc": (C.c_int, [_P, _SZ, C.POINTER(_P)]),
    "tk_device_free": (C.c_int, [_P, _P]),
    "tk_copy_to_device": (C.c_int, [_P, _P, _P, _SZ]),
    "tk_copy_to_host": (C.c_int, [_P, _P, _P, _SZ]),
    "tk_copy_device": (C.c_int, [_P, _P, _P, _SZ]),
    "tk_frame_create": (C.c_int, [_P, C.POINTER(Scene), C.POINTER(_P)]),
    "tk_frame_update": (C.c_int, [_P, C.POINTER(Scene)]),
    "tk_frame_destroy": (None, [_P]),
    "tk_frame_render": (C.c_int, [_P]),
    "tk_frame_render_tiles": (C.c_int, [_P, _P, _U32]),
    "tk_frame_set_graph": (C.c_int, [_P, C.c_int]),
    "tk_canvas_push_polygons": (C.c_int, [_P, _P, _P, C.c_uint32, _P, _P, C.POINTER(C.c_uint32)]),
    "tk_canvas_push_stroke_styled": (C.c_int, [_P, _P, C.c_size_t, _P, C.c_size_t, C.POINTER(StrokeStyle), _P, _P,
                                               C.c_uint32, C.c_double, C.POINTER(C.c_uint32)]),
    "tk_frame_set_polygons": (C.c_int, [_P, _P, _P, C.c_uint64, _P]),
    "tk_frame_set_cubic_paths": (C.c_int, [_P, _P, _P, C.c_double]),
    "tk_frame_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "tk_frame_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "tk_frame_autotune": (C.c_int, [_P, C.c_uint32, C.POINTER(Tuning)]),
    "tk_frame_render_diff": (C.c_int, [_P, _P, _P]),
    "tk_frame_backward": (C.c_int, [_P, _P, _P, _P, _P]),
    "tk_frame_diff_status": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "tk_batch_create": (C.c_int, [_P, _P, _U32, _P]),
    "tk_batch_render": (C.c_int, [_P]),
    "tk_batch_read_image": (C.c_int, [_P, _U32, _P, _SZ]),
    "tk_batch_frame": (_P, [_P]),
    "tk_batch_destroy": (None, [_P]),
    "tk_frame_read_image": (C.c_int, [_P, _P, _SZ]),
    "tk_frame_submit": (C.c_int, [_P, _P]),
    "tk_submission_is_complete": (C.c_int, [_P, _P]),
    "tk_submission_readback": (C.c_int, [_P, _P, _SZ]),
    "tk_submission_destroy": (None, [_P]),
    "tk_frame_image_device": (_P, [_P]),
    "tk_frame_read_buffer": (C.c_int, [_P, C.c_int, _P, _SZ]),
    "tk_frame_stats_read": (C.c_int, [_P, C.POINTER(FrameStats)]),
    "tk_frame_set_capture": (C.c_int, [_P, C.c_int]),
    "tk_frame_enable_timing": (C.c_int, [_P, C.c_int]),
    "tk_frame_stage_times": (C.c_int, [_P, C.POINTER(StageTimes)]),
    "tk_render_to_image": (C.c_int, [_P, C.POINTER(Scene), _P, _SZ]),
    "tk_scan_clear": (C.c_int, [_P, C.POINTER(ScanConfig), _P, _P, _P, _P, _P, _P, _P]),
    "tk_scan_count": (C.c_int, [_P, C.POINTER(ScanConfig), _P, _P, _U32, _P, _P]),
    "tk_scan_prefix": (C.c_int, [_P, C.POINTER(ScanConfig), _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "tk_scan_emit": (C.c_int, [_P, C.POINTER(ScanConfig), _P, _P, _U32, _P, _P]),
    "tk_cumsum": (C.c_int, [_P, C.POINTER(CumsumConfig), _P, _P, _P, _P, _P, _P, _P]),
    "tk_coarse": (C.c_int, [_P, C.POINTER(CoarseConfig), C.POINTER(CoarseInputs), C.POINTER(CoarseOutputs)]),
    "tk_coarse_no_alloc": (C.c_int, [_P, C.POINTER(CoarseConfig), C.POINTER(CoarseInputs), C.POINTER(CoarseOutputs)]),
    "tk_fine": (C.c_int, [_P, C.POINTER(FineConfig), C.POINTER(FineInputs), _P]),
    "tk_fill_coverage_words": (C.c_int, [_P, _U32, _P, _P, _P, _P]),
    "tk_pixel_math_words": (C.c_int, [_P, _U32, _P, _P]),
    "tk_brush_words": (C.c_int, [_P, _P, _U32, _U32, _U32, _P, _P]),
}

_lib = None


def lib():
    """Loads libtileink_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TileinkError(TK_ERR_UNAVAILABLE,
                               f"{LIB_PATH} is missing; build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        # torch's ROCm wheel carries its own libamdhip64.so.7 under the same soname.
        # Whichever copy loads first serves the whole process, and torch does not
        # find the GPU through another release's runtime, so torch (when present)
        # loads first and the library binds to its copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = C.CDLL(LIB_PATH)
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


def check(status):
    if status != TK_OK:
        raise TileinkError(status, lib().tk_last_error().decode(errors="replace"))


def ptr(array):
    """Address of a numpy array's data (None for None)."""
    if array is None:
        return None
    return array.ctypes.data


_DTYPES = {}


def view(address, dtype, count):
    """Copies `count` elements of `dtype` from a host address into a new array."""
    if dtype not in _DTYPES:
        _DTYPES[dtype] = np.dtype(dtype)  # (a ctypes structure's dtype is slow to build)
    dtype = _DTYPES[dtype]
    if count == 0 or not address:
        return np.zeros(0, dtype)
    raw = (C.c_uint8 * (dtype.itemsize * count)).from_address(address)
    return np.frombuffer(raw, dtype=dtype, count=count).copy()
