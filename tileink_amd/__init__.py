"""tileink_amd — MI355X (gfx950) backend for Tileink's compute-raster path.

Python mirror of the reference's public API for this path:

* ``Canvas`` records paths, strokes, rounded rectangles and clip layers
  (src/canvas/primitives/paths.rs, rectangles.rs, src/canvas/layers.rs);
* ``NativeRenderer(NativeBackend.Hip, width, height).render_to_image(canvas)``
  returns a submission whose ``readback()`` gives the premultiplied RGBA8 image
  (src/native/renderer.rs:177-188, src/native/submission.rs:120-146).

All work goes through libtileink_hip.so (C ABI in include/tileink_hip.h).
"""
import ctypes as C
import enum

import numpy as np

from . import _abi
from ._abi import (BRUSH_FOUR_CORNER, BRUSH_LINEAR, BRUSH_RADIAL, BRUSH_SOLID, BRUSH_SWEEP, PATTERN_BILINEAR,
                   PATTERN_NEAREST, FILL_EVEN_ODD, FILL_NONZERO, INVALID, LINE_DTYPE, PTCL_DTYPE,
                   SEGMENT_DTYPE, TileinkError, check, lib, ptr)

__all__ = [
    "Batch", "BezPath", "Canvas", "Color", "LinearGradient", "RadialGradient", "SweepGradient", "FourCornerGradient",
    "Image", "ImageBrush", "PATTERN_NEAREST", "PATTERN_BILINEAR",
    "NativeBackend", "NativeContext", "NativeRenderer", "Frame", "Join", "Cap", "FillRule",
    "SceneArrays", "TileinkError", "FILL_NONZERO", "FILL_EVEN_ODD",
]


class FillRule(enum.IntEnum):
    NonZero = FILL_NONZERO
    EvenOdd = FILL_EVEN_ODD


class Join(enum.IntEnum):
    """kurbo::Join."""
    Bevel = 0
    Miter = 1
    Round = 2


class Cap(enum.IntEnum):
    """kurbo::Cap."""
    Butt = 0
    Square = 1
    Round = 2


class NativeBackend(enum.Enum):
    """Backend selector (src/native.rs:6-11); this package provides ``Hip``."""
    Hip = "hip"


# ---------------------------------------------------------------------------
# Geometry and paint
# ---------------------------------------------------------------------------

class BezPath:
    """Move/line/quad/cubic/close path, kurbo::BezPath element order."""

    def __init__(self):
        self.verbs = []
        self.points = []

    def move_to(self, x, y):
        self.verbs.append(_abi.PATH_MOVE)
        self.points.append((x, y))
        return self

    def line_to(self, x, y):
        self.verbs.append(_abi.PATH_LINE)
        self.points.append((x, y))
        return self

    def quad_to(self, x1, y1, x2, y2):
        self.verbs.append(_abi.PATH_QUAD)
        self.points += [(x1, y1), (x2, y2)]
        return self

    def curve_to(self, x1, y1, x2, y2, x3, y3):
        self.verbs.append(_abi.PATH_CUBIC)
        self.points += [(x1, y1), (x2, y2), (x3, y3)]
        return self

    def close_path(self):
        self.verbs.append(_abi.PATH_CLOSE)
        return self

    def arrays(self):
        verbs = np.asarray(self.verbs, dtype=np.uint8)
        points = np.asarray(self.points, dtype=np.float64).reshape(-1, 2)
        return np.ascontiguousarray(verbs), np.ascontiguousarray(points)


class Color:
    """Straight-alpha sRGB colour in [0, 1] (peniko::Color)."""

    def __init__(self, r, g, b, a=1.0):
        self.rgba = (float(r), float(g), float(b), float(a))

    def _brush(self):
        b = _abi.Brush()
        b.kind = BRUSH_SOLID
        b.rgba[:] = self.rgba
        return b, None


def _gradient_brush(kind, stops, extend):
    b = _abi.Brush()
    b.kind = kind
    b.extend = int(extend)
    arr = (_abi.GradientStop * max(len(stops), 1))()
    for i, (offset, color) in enumerate(stops):
        arr[i].offset = float(offset)
        arr[i].rgba[:] = color.rgba
    b.stops = C.cast(arr, C.POINTER(_abi.GradientStop))
    b.stop_count = len(stops)
    return b, arr


def _gradient_options(b, transform, unpremultiplied):
    if transform is not None:
        b.gradient_transform[:] = tuple(float(v) for v in transform)
        b.gradient_flags |= _abi.GRADIENT_TRANSFORM
    if unpremultiplied:
        b.gradient_flags |= _abi.GRADIENT_UNPREMULTIPLIED


class LinearGradient:
    """Linear gradient brush (peniko Gradient::new_linear): stops are (offset, Color);
    extend 0 pad, 1 repeat, 2 reflect.  transform (a, b, c, d, e, f) takes user space
    to the gradient's space (the brush transform; identity when None);
    unpremultiplied interpolates the stops in straight alpha, as SVG does."""

    def __init__(self, start, end, stops, extend=0, transform=None, unpremultiplied=False):
        self.start, self.end, self.stops, self.extend = start, end, list(stops), int(extend)
        self.transform, self.unpremultiplied = transform, bool(unpremultiplied)

    def _brush(self):
        b, arr = _gradient_brush(BRUSH_LINEAR, self.stops, self.extend)
        b.start[:] = self.start
        b.end[:] = self.end
        _gradient_options(b, self.transform, self.unpremultiplied)
        return b, arr


class RadialGradient:
    """Two-point conical gradient (peniko Gradient::new_two_point_radial): circles
    (start_center, start_radius) to (end_center, end_radius); transform and
    unpremultiplied as for LinearGradient."""

    def __init__(self, start_center, start_radius, end_center, end_radius, stops, extend=0, transform=None,
                 unpremultiplied=False):
        self.start_center, self.end_center = start_center, end_center
        self.start_radius, self.end_radius = float(start_radius), float(end_radius)
        self.stops, self.extend = list(stops), int(extend)
        self.transform, self.unpremultiplied = transform, bool(unpremultiplied)

    def _brush(self):
        b, arr = _gradient_brush(BRUSH_RADIAL, self.stops, self.extend)
        _gradient_options(b, self.transform, self.unpremultiplied)
        b.start[:] = self.start_center
        b.end[:] = self.end_center
        b.start_radius, b.end_radius = self.start_radius, self.end_radius
        return b, arr


class SweepGradient:
    """Sweep (conic) gradient around `center` from start_angle to end_angle, in
    radians (peniko Gradient::new_sweep)."""

    def __init__(self, center, start_angle, end_angle, stops, extend=0):
        self.center, self.start_angle, self.end_angle = center, float(start_angle), float(end_angle)
        self.stops, self.extend = list(stops), int(extend)

    def _brush(self):
        b, arr = _gradient_brush(BRUSH_SWEEP, self.stops, self.extend)
        b.start[:] = self.center
        b.start_angle, b.end_angle = self.start_angle, self.end_angle
        return b, arr


class FourCornerGradient:
    """Bilinear blend of four corner colours over `bounds` (x0, y0, x1, y1)
    (Brush::four_corner); colours in the order top-left, top-right, bottom-right,
    bottom-left."""

    def __init__(self, bounds, colors):
        self.bounds, self.colors = tuple(float(v) for v in bounds), list(colors)

    def _brush(self):
        b = _abi.Brush()
        b.kind = BRUSH_FOUR_CORNER
        b.bounds[:] = self.bounds
        for i, color in enumerate(self.colors):
            b.corners[i][:] = color.rgba
        return b, None


def _mul_div255(a, b):
    """mul_div255 (src/shared/pixel.rs): a * b / 255 rounded, on integer arrays."""
    product = a.astype(np.uint32) * b.astype(np.uint32) + 128
    return (product + (product >> 8)) >> 8


class Image:
    """Premultiplied RGBA8 pixels (src/shared/image.rs:9-19): `pixels` is a
    (height, width) uint32 array, red in the low byte."""

    def __init__(self, pixels):
        self.pixels = np.ascontiguousarray(np.asarray(pixels, dtype=np.uint32))
        if self.pixels.ndim != 2 or self.pixels.size == 0:
            raise TileinkError(_abi.TK_ERR_INVALID, "an image is a non-empty (height, width) array")
        self.height, self.width = self.pixels.shape

    @classmethod
    def from_rgba8(cls, rgba):
        """Image::from_rgba8: straight-alpha (height, width, 4) uint8, premultiplied
        channel by channel with mul_div255 (premul_rgba8_pack, src/shared/image.rs:180-187)."""
        px = np.asarray(rgba, dtype=np.uint8)
        a = px[..., 3]
        r, g, b = (_mul_div255(px[..., k], a) for k in range(3))
        packed = r | (g << 8) | (b << 16) | (a.astype(np.uint32) << 24)
        return cls(packed.astype(np.uint32))


class ImageBrush:
    """An image of a canvas (Canvas.add_image) as a brush (Brush::Pattern with an
    image resource): `transform` (a, b, c, d, e, f) takes user space to the image's
    unit square; extend 0 pad, 1 repeat, 2 reflect; sampling PATTERN_NEAREST or
    PATTERN_BILINEAR; opacity 0-255."""

    def __init__(self, image_id, transform, extend=0, sampling=1, opacity=255):
        self.image_id, self.transform = int(image_id), tuple(float(v) for v in transform)
        self.extend, self.sampling, self.opacity = int(extend), int(sampling), int(opacity)

    def _brush(self):
        b = _abi.Brush()
        b.kind = _abi.BRUSH_IMAGE
        b.image = self.image_id
        b.pattern_transform[:] = self.transform
        b.extend, b.sampling, b.opacity = self.extend, self.sampling, self.opacity
        return b, None


def _brush_arg(brush):
    if isinstance(brush, tuple):
        brush = Color(*brush)
    b, keep = brush._brush()
    return b, keep


def _affine_arg(transform):
    if transform is None:
        return None, None
    arr = np.ascontiguousarray(np.asarray(transform, dtype=np.float64).reshape(6))
    return ptr(arr), arr


# ---------------------------------------------------------------------------
# Canvas
# ---------------------------------------------------------------------------

class SceneArrays:
    """Host copies of every array of an encoded scene (what the GPU stages read)."""

    FIELDS = ("lines", "paths", "draws", "paint", "draw_batch_ids", "layers", "batches", "tile_draws",
              "tile_draw_indices", "scan_chunks", "scan_chunk_ranges", "cumsum_chunk_offsets", "cumsum_chunk_lens",
              "cumsum_row_starts", "cumsum_row_ends", "atlas")

    def __init__(self, scene):
        s = scene
        T = s.tiles_width * s.tiles_height
        self.scalars = {name: getattr(s, name) for name, kind in _abi.Scene._fields_ if kind is not C.c_void_p}
        self.lines = _abi.view(s.lines, LINE_DTYPE, s.line_count)
        self.paths = _abi.view(s.paths, np.uint32, s.path_count * 19).reshape(-1, 19)
        self.draws = _abi.view(s.draws, np.uint32, s.draw_count * 31).reshape(-1, 31)
        self.paint = _abi.view(s.paint, np.uint32, s.paint_words)
        self.draw_batch_ids = _abi.view(s.draw_batch_ids, np.uint32, s.draw_count)
        self.layers = _abi.view(s.layers, np.uint32, s.layer_count * 3).reshape(-1, 3)
        self.batches = _abi.view(s.batches, np.uint32, s.batch_count * 2).reshape(-1, 2)
        self.tile_draws = _abi.view(s.tile_draws, np.uint32, T * 2).reshape(-1, 2)
        self.tile_draw_indices = _abi.view(s.tile_draw_indices, np.uint32, s.tile_draw_index_count)
        self.scan_chunks = _abi.view(s.scan_chunks, np.uint32, s.scan_chunk_count * 4).reshape(-1, 4)
        self.scan_chunk_ranges = _abi.view(s.scan_chunk_ranges, np.uint32, s.path_count * 2).reshape(-1, 2)
        self.cumsum_chunk_offsets = _abi.view(s.cumsum_chunk_offsets, np.uint32, s.cumsum_chunk_count)
        self.cumsum_chunk_lens = _abi.view(s.cumsum_chunk_lens, np.uint32, s.cumsum_chunk_count)
        self.cumsum_row_starts = _abi.view(s.cumsum_row_starts, np.uint32, s.cumsum_row_count)
        self.cumsum_row_ends = _abi.view(s.cumsum_row_ends, np.uint32, s.cumsum_row_count)
        self.atlas = _abi.view(s.atlas, np.uint32, s.atlas_width * s.atlas_height * s.atlas_pages) \
            if s.atlas else np.zeros(0, np.uint32)

    def __getattr__(self, name):
        scalars = self.__dict__.get("scalars", {})
        if name in scalars:
            return scalars[name]
        raise AttributeError(name)

    def to_scene(self):
        """A tk_scene pointing at these arrays (kept alive by self)."""
        s = _abi.Scene()
        for name, value in self.scalars.items():
            setattr(s, name, value)
        for name in self.FIELDS:
            arr = np.ascontiguousarray(getattr(self, name))
            setattr(self, name, arr)
            setattr(s, name, ptr(arr) if arr.size else None)
        return s


class Canvas:
    """Scene recorder (src/canvas.rs).  Coordinates are logical; scale maps to physical pixels."""

    def __init__(self, width, height, scale=1.0):
        h = C.c_void_p()
        check(lib().tk_canvas_create(int(width), int(height), float(scale), C.byref(h)))
        self._h = h
        self.width, self.height, self.scale = width, height, scale
        self._scene = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().tk_canvas_destroy(h)
            self._h = None

    @staticmethod
    def _path_args(path):
        verbs, points = path.arrays()
        return verbs, points

    def push_path(self, path, brush, transform=None, rule=FillRule.NonZero, tolerance=0.25):
        verbs, points = self._path_args(path)
        b, keep = _brush_arg(brush)
        tp, tarr = _affine_arg(transform)
        out = C.c_uint32()
        check(lib().tk_canvas_push_path(self._h, ptr(verbs), len(verbs), ptr(points), len(points), C.byref(b), tp,
                                        int(rule), float(tolerance), C.byref(out)))
        self._scene = None
        return out.value

    def push_text_outline(self, path, brush, transform=None, tolerance=0.25):
        """The outlines of a laid-out text run as one nonzero path with text
        compositing (Canvas::push_text_layout_as_path after shaping,
        src/canvas/primitives/text.rs:104-146); None for an empty outline."""
        verbs, points = self._path_args(path)
        b, keep = _brush_arg(brush)
        tp, tarr = _affine_arg(transform)
        out = C.c_uint32()
        check(lib().tk_canvas_push_text_outline(self._h, ptr(verbs), len(verbs), ptr(points), len(points), C.byref(b),
                                                tp, float(tolerance), C.byref(out)))
        self._scene = None
        return None if out.value == _abi.INVALID else out.value

    def push_polygon(self, points, brush, rule=FillRule.NonZero):
        """A closed polygon, move_to + line_to per point + close_path, from an (n, 2) array."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 2))
        if len(pts) == 0:
            return None
        verbs = np.full(len(pts) + 1, _abi.PATH_LINE, np.uint8)
        verbs[0], verbs[-1] = _abi.PATH_MOVE, _abi.PATH_CLOSE
        b, keep = _brush_arg(brush)
        out = C.c_uint32()
        check(lib().tk_canvas_push_path(self._h, ptr(verbs), len(verbs), ptr(pts), len(pts), C.byref(b), None,
                                        int(rule), 0.25, C.byref(out)))
        self._scene = None
        return out.value

    def push_polygons(self, points, counts, rgba, rules):
        """Closed polygons with solid colours in one call (tk_canvas_push_polygons):
        points (V, 2), counts (n,) points per polygon, rgba (n, 4) straight alpha,
        rules (n,).  Returns the draw ids."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 2))
        cnt = np.ascontiguousarray(np.asarray(counts, dtype=np.uint32).reshape(-1))
        col = np.ascontiguousarray(np.asarray(rgba, dtype=np.float32).reshape(-1, 4))
        rul = np.ascontiguousarray(np.asarray(rules, dtype=np.uint32).reshape(-1))
        if not (len(cnt) == len(col) == len(rul)) or int(cnt.sum()) != len(pts):
            raise TileinkError(_abi.TK_ERR_INVALID, "points, counts, colours and rules disagree")
        first = C.c_uint32()
        check(lib().tk_canvas_push_polygons(self._h, ptr(pts), ptr(cnt), len(cnt), ptr(col), ptr(rul), C.byref(first)))
        self._scene = None
        return np.arange(first.value, first.value + len(cnt), dtype=np.int64)

    def push_stroke(self, path, width, brush, transform=None, rule=FillRule.NonZero, tolerance=0.25, join=None,
                    start_cap=None, end_cap=None, miter_limit=4.0):
        """Canvas::push_stroke.  join / start_cap / end_cap (Join, Cap) default to round,
        as kurbo::Stroke::new; any other style goes through tk_canvas_push_stroke_styled."""
        verbs, points = self._path_args(path)
        b, keep = _brush_arg(brush)
        tp, tarr = _affine_arg(transform)
        out = C.c_uint32()
        if join is None and start_cap is None and end_cap is None:
            check(lib().tk_canvas_push_stroke(self._h, ptr(verbs), len(verbs), ptr(points), len(points), float(width),
                                              C.byref(b), tp, int(rule), float(tolerance), C.byref(out)))
        else:
            st = _abi.StrokeStyle(float(width), float(miter_limit), int(Join.Round if join is None else join),
                                  int(Cap.Round if start_cap is None else start_cap),
                                  int(Cap.Round if end_cap is None else end_cap), 0)
            check(lib().tk_canvas_push_stroke_styled(self._h, ptr(verbs), len(verbs), ptr(points), len(points),
                                                     C.byref(st), C.byref(b), tp, int(rule), float(tolerance),
                                                     C.byref(out)))
        self._scene = None
        return out.value

    def push_rect(self, x0, y0, x1, y1, radii, brush):
        r = (C.c_float * 4)(*((radii,) * 4 if np.isscalar(radii) else radii))
        b, keep = _brush_arg(brush)
        out = C.c_uint32()
        check(lib().tk_canvas_push_rect(self._h, float(x0), float(y0), float(x1), float(y1), r, C.byref(b),
                                        C.byref(out)))
        self._scene = None
        return out.value

    def add_image(self, image):
        """Registers an Image (Canvas::register_scene_image); returns its id for
        ImageBrush and push_image."""
        if not isinstance(image, Image):
            image = Image(image)
        out = C.c_uint32()
        check(lib().tk_canvas_add_image(self._h, image.width, image.height, ptr(image.pixels), C.byref(out)))
        self._scene = None
        return out.value

    def push_image(self, x0, y0, x1, y1, image_id, extend=0, sampling=PATTERN_BILINEAR):
        """Canvas::push_image: the image scaled into the rect as an SDF rect draw;
        None when the rect is empty or not finite."""
        out = C.c_uint32()
        check(lib().tk_canvas_push_image(self._h, float(x0), float(y0), float(x1), float(y1), int(image_id),
                                         int(extend), int(sampling), C.byref(out)))
        self._scene = None
        return None if out.value == _abi.INVALID else out.value

    def push_clip_sdf_rect_layer(self, x0, y0, x1, y1, radii=0.0):
        r = (C.c_float * 4)(*((radii,) * 4 if np.isscalar(radii) else radii))
        check(lib().tk_canvas_push_clip_sdf_rect_layer(self._h, float(x0), float(y0), float(x1), float(y1), r))
        self._scene = None

    def push_clip_layer(self, path, transform=None, rule=FillRule.NonZero, tolerance=0.25):
        verbs, points = self._path_args(path)
        tp, tarr = _affine_arg(transform)
        check(lib().tk_canvas_push_clip_layer(self._h, ptr(verbs), len(verbs), ptr(points), len(points), tp, int(rule),
                                              float(tolerance)))
        self._scene = None

    def pop_layer(self):
        check(lib().tk_canvas_pop_layer(self._h))
        self._scene = None

    def line_count(self):
        return lib().tk_canvas_line_count(self._h)

    def encode(self, clear_color=0):
        """Compiles and plans the scene; returns the tk_scene view (owned by the canvas)."""
        s = _abi.Scene()
        check(lib().tk_canvas_encode(self._h, int(clear_color), C.byref(s)))
        s._canvas = self  # the scene points into this canvas's arrays
        self._scene = s
        return s

    def arrays(self, clear_color=0):
        return SceneArrays(self.encode(clear_color))


# ---------------------------------------------------------------------------
# Device
# ---------------------------------------------------------------------------

class NativeContext:
    """One gfx950 device, chosen explicitly (NativeContextOptions::physical_adapter)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().tk_context_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().tk_context_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        check(lib().tk_context_synchronize(self._h))

    def stream_handle(self):
        """The context's HIP stream (hipStream_t as an integer): every call on this
        context is ordered on it."""
        return int(lib().tk_context_stream(self._h) or 0)

    def to_device(self, array):
        """Copies a numpy array into a new device buffer; returns its DeviceBuffer."""
        array = np.ascontiguousarray(array)
        buf = DeviceBuffer(self, max(array.nbytes, 4))
        check(lib().tk_copy_to_device(self._h, buf.ptr, ptr(array), array.nbytes))
        return buf

    def limits(self):
        out = _abi.Limits()
        check(lib().tk_context_limits(self._h, C.byref(out)))
        return {f: getattr(out, f) for f, _ in out._fields_}


class DeviceBuffer:
    """Device allocation owned by a NativeContext."""

    def __init__(self, context, nbytes):
        self.context = context
        self.nbytes = int(nbytes)
        h = C.c_void_p()
        check(lib().tk_device_alloc(context.handle, self.nbytes, C.byref(h)))
        self._ptr = h

    @property
    def ptr(self):
        return self._ptr.value

    def read(self, dtype, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        check(lib().tk_copy_to_host(self.context.handle, ptr(out), self.ptr, out.nbytes))
        return out

    def write(self, array):
        array = np.ascontiguousarray(array)
        if array.nbytes > self.nbytes:
            raise TileinkError(_abi.TK_ERR_INVALID, "write larger than the buffer")
        check(lib().tk_copy_to_device(self.context.handle, self.ptr, ptr(array), array.nbytes))

    def __del__(self):
        p = getattr(self, "_ptr", None)
        if p is not None and p.value:
            lib().tk_device_free(self.context.handle, p)
            self._ptr = None


class Frame:
    """Device-resident scene: inputs uploaded once, rendered any number of times."""

    def __init__(self, context, scene):
        self.context = context
        self._scene = scene
        h = C.c_void_p()
        check(lib().tk_frame_create(context.handle, C.byref(scene), C.byref(h)))
        self._h = h
        self.width, self.height = scene.width, scene.height
        self.batch_count = scene.batch_count
        self.tiles = scene.tiles_width * scene.tiles_height

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and not getattr(self, "_borrowed", False):
            lib().tk_frame_destroy(h)
            self._h = None

    def render(self):
        check(lib().tk_frame_render(self._h))

    def render_tiles(self, tiles):
        """Incremental render of the listed tiles (row-major indices); other tiles keep
        their pixels (tk_frame_render_tiles)."""
        tiles = np.ascontiguousarray(np.asarray(tiles, dtype=np.uint32).reshape(-1))
        check(lib().tk_frame_render_tiles(self._h, ptr(tiles) if tiles.size else None, int(tiles.size)))

    def set_polygons(self, points, starts, line_vertex=False):
        """GPU front end (tk_frame_set_polygons): new vertices for every path of a
        polygon scene; the device rebuilds lines, path records, bounds, tile bins
        and plans.  points: (N, 2) float32 and starts: (path_count + 1,) uint32,
        either torch tensors on the GPU (read in place) or numpy arrays (uploaded).
        With line_vertex, returns each line's first-vertex index (a torch int32
        tensor for torch input, else a numpy array), trimmed to the line count."""
        keep = []

        def device(arr, dtype, torch_dtype):
            if hasattr(arr, "data_ptr"):
                t = arr.contiguous()
                if t.dtype != torch_dtype or not t.is_cuda:
                    raise TileinkError(_abi.TK_ERR_INVALID, f"expected a {torch_dtype} tensor on the GPU")
                keep.append(t)
                return t.data_ptr(), t.numel()
            a = np.ascontiguousarray(arr, dtype=dtype).reshape(-1)
            buf = self.context.to_device(a if a.size else np.zeros(1, dtype))
            keep.append(buf)
            return buf.ptr, a.size

        import importlib
        torch = importlib.import_module("torch") if any(hasattr(x, "data_ptr") for x in (points, starts)) else None
        pp, n = device(points, np.float32, getattr(torch, "float32", None))
        sp, ns = device(starts, np.uint32, getattr(torch, "int32", None))
        if n % 2 or ns != self._path_count() + 1:
            raise TileinkError(_abi.TK_ERR_INVALID, "points are (x, y) pairs and starts has path_count + 1 entries")
        out = None
        if line_vertex:
            if torch is not None:
                out = torch.empty(max(n // 2, 1), dtype=torch.int32, device=keep[0].device)
                optr = out.data_ptr()
            else:
                out = DeviceBuffer(self.context, 4 * max(n // 2, 1))
                optr = out.ptr
        check(lib().tk_frame_set_polygons(self._h, C.c_void_p(pp), C.c_void_p(sp), n // 2,
                                          C.c_void_p(optr) if line_vertex else None))
        if not line_vertex:
            return None
        lines = int(self.stats()["lines"])
        return out[:lines] if torch is not None else out.read(np.uint32, lines)

    def set_cubic_paths(self, ctrl, starts, tolerance=0.25):
        """GPU flattening + front end (tk_frame_set_cubic_paths) for scenes whose
        paths are each one closed contour of cubics: ctrl (N, 2) float64 control
        points (start point, then three per cubic, path after path) and starts
        (path_count + 1,), numpy arrays (uploaded) or torch tensors on the GPU."""
        keep = []

        def device(arr, dtype):
            if hasattr(arr, "data_ptr"):
                t = arr.contiguous()
                keep.append(t)
                return t.data_ptr(), t.numel()
            a = np.ascontiguousarray(arr, dtype=dtype).reshape(-1)
            buf = self.context.to_device(a if a.size else np.zeros(1, dtype))
            keep.append(buf)
            return buf.ptr, a.size

        cp, n = device(ctrl, np.float64)
        sp, ns = device(starts, np.uint32)
        if n % 2 or ns != self._path_count() + 1:
            raise TileinkError(_abi.TK_ERR_INVALID, "ctrl holds (x, y) pairs and starts has path_count + 1 entries")
        check(lib().tk_frame_set_cubic_paths(self._h, C.c_void_p(cp), C.c_void_p(sp), float(tolerance)))

    def _path_count(self):
        return int(self.stats()["paths"])

    def update(self, scene):
        """Replaces the scene's arrays; every size must match (tk_frame_update)."""
        check(lib().tk_frame_update(self._h, C.byref(scene)))
        self._scene = scene

    def tuning(self):
        """The launch parameters as a dict (tk_frame_get_tuning)."""
        t = _abi.Tuning()
        check(lib().tk_frame_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k in _abi.TUNING_FIELDS}

    def set_tuning(self, scan_xcd_run=None, coarse_xcd_run=None, fine_xcd_run=None, **more):
        """Sets launch parameters (tk_frame_set_tuning); those not given keep their
        values.  Keywords: the tk_tuning fields (persistent, scan_block,
        coarse_block, coarse_scan_items and the three XCD run lengths)."""
        cur = self.tuning()
        given = dict(scan_xcd_run=scan_xcd_run, coarse_xcd_run=coarse_xcd_run, fine_xcd_run=fine_xcd_run, **more)
        unknown = set(given) - set(_abi.TUNING_FIELDS)
        if unknown:
            raise TileinkError(_abi.TK_ERR_INVALID, f"unknown tuning fields {sorted(unknown)}")
        t = _abi.Tuning(*(cur[k] if given.get(k) is None else int(given[k]) for k in _abi.TUNING_FIELDS))
        check(lib().tk_frame_set_tuning(self._h, C.byref(t)))

    def autotune(self, repeats=3):
        """Times candidate launch parameters on this scene and keeps the fastest (tk_frame_autotune)."""
        t = _abi.Tuning()
        check(lib().tk_frame_autotune(self._h, int(repeats), C.byref(t)))
        return {k: getattr(t, k) for k in _abi.TUNING_FIELDS}

    def set_graph(self, enable=True):
        """Replay full renders from a captured hipGraph (tk_frame_set_graph)."""
        check(lib().tk_frame_set_graph(self._h, int(bool(enable))))

    def set_capture(self, enable=True):
        check(lib().tk_frame_set_capture(self._h, int(bool(enable))))

    def enable_timing(self, enable=True):
        check(lib().tk_frame_enable_timing(self._h, int(bool(enable))))

    def stage_times(self):
        out = _abi.StageTimes()
        check(lib().tk_frame_stage_times(self._h, C.byref(out)))
        return {f: getattr(out, f) for f, _ in out._fields_}

    def image(self):
        out = np.empty((self.height, self.width), dtype=np.uint32)
        check(lib().tk_frame_read_image(self._h, ptr(out), out.size))
        return out

    def submit(self):
        """Enqueues a render and returns its Submission without waiting (tk_frame_submit)."""
        h = C.c_void_p()
        check(lib().tk_frame_submit(self._h, C.byref(h)))
        return Submission(h, self.width, self.height, self)

    def stats(self):
        out = _abi.FrameStats()
        check(lib().tk_frame_stats_read(self._h, C.byref(out)))
        return {f: getattr(out, f) for f, _ in out._fields_}

    def buffer(self, which, dtype, count):
        out = np.empty(count, dtype=dtype)
        check(lib().tk_frame_read_buffer(self._h, int(which), ptr(out), out.nbytes))
        return out


class Batch:
    """Independent scenes rendered as one frame (tk_batch_*): `scenes` are encoded
    tk_scene views (Canvas.encode / SceneArrays.to_scene), kept alive by the batch."""

    def __init__(self, context, scenes):
        self.context = context
        self._scenes = list(scenes)
        arr = (_abi.Scene * len(self._scenes))(*self._scenes)
        h = C.c_void_p()
        check(lib().tk_batch_create(context.handle, arr, len(self._scenes), C.byref(h)))
        self._h = h
        self.sizes = [(sc.width, sc.height) for sc in self._scenes]
        frame = Frame.__new__(Frame)
        frame._borrowed = True
        frame._h = C.c_void_p(lib().tk_batch_frame(h))
        frame.context, frame._scene = context, None
        self.frame = frame  # stats, stage times, timing and graph switches

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().tk_batch_destroy(h)
            self._h = None

    def __len__(self):
        return len(self._scenes)

    def render(self):
        check(lib().tk_batch_render(self._h))

    def image(self, index):
        if not 0 <= index < len(self.sizes):
            raise TileinkError(_abi.TK_ERR_INVALID, f"scene index {index} out of range ({len(self.sizes)} scenes)")
        w, h = self.sizes[index]
        out = np.empty((h, w), dtype=np.uint32)
        check(lib().tk_batch_read_image(self._h, int(index), ptr(out), out.size))
        return out


class Submission:
    """A submitted frame and its completion receipt (NativeImageSubmission,
    src/native/submission.rs:100-146): is_complete() never waits; readback()
    waits, then returns the (H, W) premultiplied RGBA8 image, raising a
    TileinkError with TK_ERR_OVERFLOW when the frame exceeded a capacity."""

    def __init__(self, handle, width, height, owner=None):
        self._h, self.width, self.height = handle, width, height
        self._owner = owner  # keeps the frame (and its context) alive while the work runs

    def is_complete(self):
        done = C.c_uint32()
        check(lib().tk_submission_is_complete(self._h, C.byref(done)))
        return bool(done.value)

    def readback(self):
        out = np.empty((self.height, self.width), dtype=np.uint32)
        check(lib().tk_submission_readback(self._h, ptr(out), out.size))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().tk_submission_destroy(h)
            self._h = None


class ImageSubmission(Submission):
    """Result of render_to_image (src/native/submission.rs:100-146): the frame is
    submitted when it is returned and read back on readback()."""


class NativeRenderer:
    """NativeRenderer::new(backend, width, height) + render_to_image (src/native/renderer.rs:51-188)."""

    def __init__(self, backend, width, height, device=0, context=None):
        if backend is not NativeBackend.Hip:
            raise TileinkError(_abi.TK_ERR_UNAVAILABLE, f"backend {backend} is not provided by this package")
        self.width, self.height = width, height
        self.context = context or NativeContext(device)

    def render_to_image(self, canvas, clear_color=0):
        scene = canvas.encode(clear_color)
        if (scene.width, scene.height) != (self.width, self.height):
            raise TileinkError(_abi.TK_ERR_INVALID,
                               f"canvas is {scene.width}x{scene.height} physical pixels, renderer {self.width}x{self.height}")
        frame = Frame(self.context, scene)
        h = C.c_void_p()
        check(lib().tk_frame_submit(frame._h, C.byref(h)))
        return ImageSubmission(h, scene.width, scene.height, frame)


def line_scanned_tile_count(p0, p1, bbox):
    """CPU tile count of a line within a tile bbox (src/shared/scan_line.rs:17-31)."""
    line = _abi.Line(0, 0.0, (C.c_float * 2)(*p0), (C.c_float * 2)(*p1))
    return lib().tk_line_scanned_tile_count(C.byref(line), *[int(v) for v in bbox])
