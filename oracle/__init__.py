"""Parity oracle: ctypes bindings of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.or_render_frame.restype = C.c_int
        _lib.or_line_scanned_tile_count.restype = C.c_uint32
        _lib.or_polygon_pixel_area_f64.restype = C.c_double
        _lib.or_polygon_pixel_area_f64.argtypes = [C.c_void_p, C.c_uint32, C.c_double, C.c_double]
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def u32(a):
    return np.ascontiguousarray(np.asarray(a).view(np.uint32) if np.asarray(a).dtype.itemsize == 4 and
                                np.asarray(a).dtype != np.uint32 else np.asarray(a, dtype=np.uint32))


class ScanConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("clear_len", "backdrop_len", "path_count", "scan_chunk_count", "line_count",
                                          "segment_capacity", "incremental", "line_base", "path_base", "chunk_base",
                                          "backdrop_base")]


class CoarseConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("tile_count", "tiles_width", "tiles_height", "batch_id",
                                          "layer_stack_start", "layer_stack_end", "ptcl_capacity", "paint_brush_base",
                                          "draw_count", "path_count", "paint_words", "layer_count",
                                          "tile_draw_index_count")]


class CoarseInputs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("draws", "paths", "backdrops", "segment_ranges", "layers", "paint",
                                          "draw_batch_ids", "tile_draws", "tile_draw_indices")]


class Scene(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("tiles_width", C.c_uint32), ("tiles_height", C.c_uint32),
        ("clear_color", C.c_uint32),
        ("lines", C.c_void_p), ("line_count", C.c_uint32),
        ("paths", C.c_void_p), ("path_count", C.c_uint32),
        ("draws", C.c_void_p), ("draw_count", C.c_uint32),
        ("paint", C.c_void_p), ("paint_words", C.c_uint32), ("paint_sdf_shadow_base", C.c_uint32),
        ("paint_brush_base", C.c_uint32),
        ("draw_batch_ids", C.c_void_p),
        ("layers", C.c_void_p), ("layer_count", C.c_uint32),
        ("batches", C.c_void_p), ("batch_count", C.c_uint32),
        ("tile_draws", C.c_void_p), ("tile_draw_indices", C.c_void_p), ("tile_draw_index_count", C.c_uint32),
        ("scan_chunks", C.c_void_p), ("scan_chunk_count", C.c_uint32),
        ("scan_chunk_ranges", C.c_void_p),
        ("cumsum_chunk_offsets", C.c_void_p), ("cumsum_chunk_lens", C.c_void_p), ("cumsum_chunk_count", C.c_uint32),
        ("cumsum_row_starts", C.c_void_p), ("cumsum_row_ends", C.c_void_p), ("cumsum_row_count", C.c_uint32),
        ("backdrop_len", C.c_uint32), ("segment_capacity", C.c_uint32), ("max_clip_depth", C.c_uint32),
        ("atlas", C.c_void_p), ("atlas_width", C.c_uint32), ("atlas_height", C.c_uint32), ("atlas_pages", C.c_uint32),
    ]


class FrameOutputs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("raw_backdrops", "backdrops", "segment_counts", "segment_ranges",
                                          "segment_cursors", "segment_bumps", "segments", "tile_ptcl_counts",
                                          "tile_ptcl_starts", "tile_ptcl_ends", "tile_kinds", "batch_bases",
                                          "ptcl")] + [("ptcl_capacity_words", C.c_uint64), ("stage_ms", C.c_void_p)]

STAGES = ("scan_count", "scan_prefix", "scan_emit", "cumsum", "coarse", "fine")


def set_threads(n):
    """Threads for the oracle's independent work items (results do not depend on it)."""
    lib().or_set_threads(C.c_uint32(int(n)))


def threads():
    lib().or_get_threads.restype = C.c_uint32
    return int(lib().or_get_threads())


ARRAY_FIELDS = ("lines", "paths", "draws", "paint", "draw_batch_ids", "layers", "batches", "tile_draws",
                "tile_draw_indices", "scan_chunks", "scan_chunk_ranges", "cumsum_chunk_offsets", "cumsum_chunk_lens",
                "cumsum_row_starts", "cumsum_row_ends", "atlas")


def _scene_struct(scene):
    """An or_scene over the arrays of `scene` (a tileink_amd.SceneArrays); the
    second value keeps the word copies alive."""
    s = Scene()
    keep = []
    for name, _ in Scene._fields_:
        if name in ARRAY_FIELDS:
            arr = np.ascontiguousarray(np.asarray(getattr(scene, name)))
            raw = np.ascontiguousarray(arr.view(np.uint32).reshape(-1)) if arr.size else np.zeros(1, np.uint32)
            keep.append(raw)
            setattr(s, name, raw.ctypes.data)
        else:
            setattr(s, name, int(scene.scalars[name]))
    return s, keep


def tile_draw_bins(scene, cull=True):
    """Tile draw lists of an immediate canvas (TileDrawBins::reset_transient), with
    early occlusion unless cull is False.  Returns (tile_draws[T, 2] u32 {start,
    len}, indices u32)."""
    s, keep = _scene_struct(scene)
    T = s.tiles_width * s.tiles_height
    tile_draws = np.zeros((T, 2), np.uint32)
    capacity = 0
    tiles_w, tiles_h = s.tiles_width, s.tiles_height
    draws = np.asarray(scene.draws, np.uint32).reshape(-1, 31)
    ids = np.asarray(scene.draw_batch_ids, np.uint32)
    for d in range(len(draws)):  # every member's tile box bounds its references
        if ids[d] >= s.batch_count:
            continue
        pb = draws[d, 10:14].view(np.int32).astype(np.int64)
        x0 = min(max(pb[0], 0) // 16, tiles_w)
        y0 = min(max(pb[1], 0) // 16, tiles_h)
        x1 = min((max(pb[2], 0) + 15) // 16, tiles_w)
        y1 = min((max(pb[3], 0) + 15) // 16, tiles_h)
        capacity += max(x1 - x0, 0) * max(y1 - y0, 0)
    indices = np.zeros(max(capacity, 1), np.uint32)
    lib().or_tile_draw_bins.restype = C.c_int64
    n = lib().or_tile_draw_bins(C.byref(s), C.c_int(1 if cull else 0), _p(tile_draws), _p(indices),
                                C.c_uint64(indices.size))
    if n < 0:
        raise RuntimeError("tile draw references exceed their bound")
    return tile_draws, indices[:n].copy()


def render_frame(scene, ptcl_capacity=None, outputs=True):
    """Runs the whole frame on the CPU.  `scene` is a tileink_amd.SceneArrays.

    Returns (image[h, w] u32, dict of intermediate arrays and the stage times in
    ms under "stage_ms"); outputs=False returns only the image and stage times."""
    s, keep = _scene_struct(scene)
    if not outputs:
        fo = FrameOutputs()
        ms = np.zeros(6, np.float64)
        fo.stage_ms = ms.ctypes.data
        image = np.zeros((s.height, s.width), np.uint32)
        status = lib().or_render_frame(C.byref(s), _p(image), C.byref(fo))
        if status != 0:
            raise RuntimeError(f"oracle frame exceeded a bound (status {status})")
        return image, {"stage_ms": dict(zip(STAGES, ms.tolist()))}
    B, P = s.backdrop_len, s.path_count
    T = s.tiles_width * s.tiles_height
    Q = s.batch_count
    if ptcl_capacity is None:
        ptcl_capacity = int(scene.scalars["ptcl_capacity"])
    out = {
        "raw_backdrops": np.zeros(B, np.int32), "backdrops": np.zeros(B, np.int32),
        "segment_counts": np.zeros(B, np.uint32), "segment_ranges": np.zeros((B, 2), np.uint32),
        "segment_cursors": np.zeros(B, np.uint32), "segment_bumps": np.zeros(P, np.uint32),
        "segments": np.zeros((s.segment_capacity, 5), np.uint32),
        "tile_ptcl_counts": np.zeros(Q * T, np.uint32), "tile_ptcl_starts": np.zeros(Q * T, np.uint32),
        "tile_ptcl_ends": np.zeros(Q * T, np.uint32), "tile_kinds": np.zeros(Q * T, np.uint32),
        "batch_bases": np.zeros(Q + 1, np.uint32), "ptcl": np.zeros((max(ptcl_capacity, 1), 6), np.uint32),
    }
    fo = FrameOutputs()
    for name, arr in out.items():
        setattr(fo, name, arr.ctypes.data)
    fo.ptcl_capacity_words = out["ptcl"].size
    ms = np.zeros(6, np.float64)
    fo.stage_ms = ms.ctypes.data
    image = np.zeros((s.height, s.width), np.uint32)
    status = lib().or_render_frame(C.byref(s), _p(image), C.byref(fo))
    if status != 0:
        raise RuntimeError(f"oracle frame exceeded a bound (status {status})")
    out["ptcl"] = out["ptcl"][: int(out["batch_bases"][Q])]
    out["stage_ms"] = dict(zip(STAGES, ms.tolist()))
    return image, out


def scan_count(cfg, lines, paths, path_record_count, backdrops, counts):
    lib().or_scan_count(C.byref(cfg), _p(lines), _p(paths), C.c_uint32(path_record_count), _p(backdrops), _p(counts))


def scan_prefix(cfg, chunks, chunk_ranges, paths, counts, ranges, cursors, bumps, totals, offsets):
    """The reference's three prefix passes in order."""
    lib().or_scan_prefix_chunks(C.byref(cfg), _p(chunks), _p(counts), _p(ranges), _p(totals))
    lib().or_scan_chunk_offsets(C.byref(cfg), _p(paths), _p(chunk_ranges), _p(bumps), _p(totals), _p(offsets))
    lib().or_scan_apply_chunk_offsets(C.byref(cfg), _p(chunks), _p(ranges), _p(cursors), _p(offsets))


def coarse_prefix(tile_count, work, chunk_records):
    lib().or_coarse_prefix(C.c_uint32(tile_count), _p(work), _p(chunk_records))


def coarse_count_bins(cfg, inputs, work, kinds):
    lib().or_coarse_count_bins(C.byref(cfg), C.byref(inputs), _p(work), _p(kinds))


def coarse_emit_bins(cfg, inputs, work, kinds):
    lib().or_coarse_emit_bins(C.byref(cfg), C.byref(inputs), _p(work), _p(kinds))


def line_scanned_tile_count(p0, p1, bbox):
    a = np.asarray(p0, np.float32)
    b = np.asarray(p1, np.float32)
    return lib().or_line_scanned_tile_count(_p(a), _p(b), *[C.c_uint32(int(v)) for v in bbox])


def fill_coverage_words(requests, segments, sequential=False):
    """Alpha and FP32 coverage per request.  Default: the three sums exact (the form
    the fine stage computes); sequential=True: FP32 sums in stored segment order."""
    requests = np.ascontiguousarray(requests, dtype=np.uint32).reshape(-1, 8)
    segments = np.ascontiguousarray(segments).view(np.uint32).reshape(-1)
    n = requests.shape[0]
    alpha = np.zeros(n, np.uint32)
    cov = np.zeros(n, np.float32)
    fn = lib().or_fill_coverage_words_sequential if sequential else lib().or_fill_coverage_words
    fn(C.c_uint32(n), _p(requests), _p(segments), _p(alpha), _p(cov))
    return alpha, cov


def fill_coverage_f64(requests, segments):
    requests = np.ascontiguousarray(requests, dtype=np.uint32).reshape(-1, 8)
    segments = np.ascontiguousarray(segments).view(np.uint32).reshape(-1)
    out = np.zeros(requests.shape[0], np.float64)
    lib().or_fill_coverage_f64(C.c_uint32(requests.shape[0]), _p(requests), _p(segments), _p(out))
    return out


def polygon_pixel_area_f64(lines, px, py):
    lines = np.ascontiguousarray(lines).view(np.uint32).reshape(-1)
    return lib().or_polygon_pixel_area_f64(_p(lines), C.c_uint32(lines.size // 6), float(px), float(py))


def pixel_math_words(requests):
    requests = np.ascontiguousarray(requests, dtype=np.uint32).reshape(-1, 8)
    out = np.zeros((requests.shape[0], 14), np.uint32)
    lib().or_pixel_math_words(C.c_uint32(requests.shape[0]), _p(requests), _p(out))
    return out


def text_words(requests):
    """Text compositing of path glyphs per request {destination, source, coverage,
    foreground f32 bits}: (n, 4) words {composite, corrected coverage bits,
    srgb_to_linear bits, linear_to_srgb bits}."""
    requests = np.ascontiguousarray(requests, dtype=np.uint32).reshape(-1, 4)
    out = np.zeros((requests.shape[0], 4), np.uint32)
    lib().or_text_words(C.c_uint32(requests.shape[0]), _p(requests), _p(out))
    return out


def brush_words(paint, paint_brush_base, requests):
    """sample_brush at each request {data_base, x_bits, y_bits, 0} (validation/brush.hlsl:14-18)."""
    paint = np.ascontiguousarray(paint, dtype=np.uint32).reshape(-1)
    requests = np.ascontiguousarray(requests, dtype=np.uint32).reshape(-1, 4)
    out = np.zeros(requests.shape[0], np.uint32)
    lib().or_brush_words(_p(paint), C.c_uint32(paint.size), C.c_uint32(paint_brush_base),
                         C.c_uint32(requests.shape[0]), _p(requests), _p(out))
    return out


# --- the differentiable form in FP64 ---------------------------------------------

def _diff_args(draws, lines):
    starts = np.ascontiguousarray([d[0] for d in draws], np.uint32)
    counts = np.ascontiguousarray([d[1] for d in draws], np.uint32)
    rules = np.ascontiguousarray([d[2] for d in draws], np.uint32)
    lines = np.ascontiguousarray(lines, np.float64).reshape(-1, 4)
    return starts, counts, rules, lines


def diff_area(lines, start, count, width, height):
    """(box (x0, y0, w, h), per-pixel winding integral over the box) of lines[start:start+count]."""
    lines = np.ascontiguousarray(lines, np.float64).reshape(-1, 4)
    box = np.zeros(4, np.int32)
    lib().or_diff_area_f64(_p(lines), C.c_uint32(start), C.c_uint32(count), C.c_uint32(width),
                           C.c_uint32(height), _p(box), None)
    area = np.zeros((box[3], box[2]), np.float64)
    lib().or_diff_area_f64(_p(lines), C.c_uint32(start), C.c_uint32(count), C.c_uint32(width),
                           C.c_uint32(height), _p(box), _p(area))
    return box, area


def diff_render(width, height, clear, draws, colors, lines, masks=None):
    """FP64 forward of the differentiable form.  draws: (line_start, line_count, rule)
    per draw in painter order; colors: premultiplied RGBA per draw; lines: (L, 4).
    masks: optional per-draw clip masks (H, W) (None where a draw has no clip);
    a masked draw's alpha is multiplied by its mask.  A colour may also be an
    (H, W, 4) image (a gradient paint evaluated at the pixel centres)."""
    if (masks is not None and any(m is not None for m in masks)) or _per_pixel(colors):
        return _masked_render(width, height, clear, draws, colors, lines, masks)
    starts, counts, rules, lines = _diff_args(draws, lines)
    clear = np.ascontiguousarray(clear, np.float64)
    colors = np.ascontiguousarray(colors, np.float64).reshape(-1, 4)
    image = np.zeros((height, width, 4), np.float64)
    lib().or_diff_render_f64(C.c_uint32(width), C.c_uint32(height), _p(clear), C.c_uint32(len(draws)), _p(starts),
                             _p(counts), _p(rules), _p(colors), _p(lines), _p(image))
    return image


def path_clip_coverage(points, rule, width, height):
    """FP64 mask (H, W) of a path clip layer: the exact area coverage of the closed
    polygon `points` (n, 2) at every pixel, turned into alpha by its fill rule (0
    NonZero, 1 EvenOdd), as the differentiable form's clip masks are."""
    p = np.asarray(points, np.float64).reshape(-1, 2)
    lines = np.concatenate([p, np.roll(p, -1, 0)], 1)
    box, area = diff_area(lines, 0, len(lines), width, height)
    full = np.zeros((height, width), np.float64)
    x0, y0, w, h = (int(b) for b in box)
    full[y0:y0 + h, x0:x0 + w] = area
    return _alpha(full, int(rule))[0]


def _per_pixel(colors):
    return isinstance(colors, (list, tuple)) and any(np.ndim(c) == 3 for c in colors)


def _alpha(a, rule):
    if rule == 1:
        w = a - 2.0 * np.rint(0.5 * a)
        return np.abs(w), np.where(w < 0, -1.0, 1.0)
    m = np.abs(a)
    return np.minimum(m, 1.0), np.where(m < 1.0, np.where(a < 0, -1.0, 1.0), 0.0)


def _masked_render(width, height, clear, draws, colors, lines, masks, image_grad=None):
    """The composite with clip masks, per draw over its box: the areas come from
    or_diff_area_f64; the composite and its adjoint are restated here in numpy."""
    lines = np.ascontiguousarray(lines, np.float64).reshape(-1, 4)
    if not _per_pixel(colors):
        colors = list(np.asarray(colors, np.float64).reshape(-1, 4))
    masks = masks if masks is not None else [None] * len(draws)
    image = np.broadcast_to(np.asarray(clear, np.float64), (height, width, 4)).copy()
    layers = []
    for d, (start, count, rule) in enumerate(draws):
        box, area = diff_area(lines, start, count, width, height)
        x0, y0, w, h = (int(v) for v in box)
        alpha, slope = _alpha(area, rule)
        if masks[d] is not None:
            m = np.asarray(masks[d], np.float64)[y0:y0 + h, x0:x0 + w]
            alpha, slope = alpha * m, slope * m
        c = np.asarray(colors[d], np.float64)
        c = c[y0:y0 + h, x0:x0 + w] if c.ndim == 3 else np.broadcast_to(c, (h, w, 4))
        layers.append((x0, y0, w, h, alpha, slope, c))
    if image_grad is None:
        for x0, y0, w, h, alpha, _, c in layers:
            px = image[y0:y0 + h, x0:x0 + w]
            px *= (1.0 - alpha * c[..., 3])[..., None]
            px += alpha[..., None] * c
        return image
    g = np.asarray(image_grad, np.float64).reshape(height, width, 4)
    T = np.ones((height, width))
    trans = [None] * len(layers)
    for d in range(len(layers) - 1, -1, -1):
        x0, y0, w, h, alpha, _, c = layers[d]
        trans[d] = T[y0:y0 + h, x0:x0 + w].copy()
        T[y0:y0 + h, x0:x0 + w] *= 1.0 - alpha * c[..., 3]
    color_grad, color_maps = np.zeros((len(layers), 4)), []
    cov_grad = []
    for d, (x0, y0, w, h, alpha, slope, c) in enumerate(layers):
        r = image[y0:y0 + h, x0:x0 + w]
        gg = g[y0:y0 + h, x0:x0 + w]
        t = trans[d]
        g_r = (gg * r).sum(-1)
        g_c = (gg * c).sum(-1)
        cov_grad.append((t * (g_c - c[..., 3] * g_r) * slope, x0, y0, w, h))
        cmap = np.concatenate([(t * alpha)[..., None] * gg[..., :3], (t * alpha * (gg[..., 3] - g_r))[..., None]], -1)
        color_grad[d] = cmap.sum((0, 1))
        full = np.zeros((height, width, 4))
        full[y0:y0 + h, x0:x0 + w] = cmap
        color_maps.append(full)
        r *= (1.0 - alpha * c[..., 3])[..., None]
        r += alpha[..., None] * c
    return cov_grad, (color_grad, color_maps)


def diff_backward(width, height, clear, draws, colors, lines, image_grad, masks=None, color_maps=False):
    """FP64 analytic gradient: (dL/d lines (L, 4), dL/d colours (D, 4)) for dL/d image;
    with color_maps=True the second item is the per-pixel dL/d colour (H, W, 4) of
    each draw instead (for per-pixel colours)."""
    if (masks is not None and any(m is not None for m in masks)) or _per_pixel(colors) or color_maps:
        return _masked_backward(width, height, clear, draws, colors, lines, image_grad, masks, color_maps)
    starts, counts, rules, lines = _diff_args(draws, lines)
    clear = np.ascontiguousarray(clear, np.float64)
    colors = np.ascontiguousarray(colors, np.float64).reshape(-1, 4)
    image_grad = np.ascontiguousarray(image_grad, np.float64).reshape(height, width, 4)
    line_grad = np.zeros_like(lines)
    color_grad = np.zeros((len(draws), 4), np.float64)
    lib().or_diff_backward_f64(C.c_uint32(width), C.c_uint32(height), _p(clear), C.c_uint32(len(draws)),
                               _p(starts), _p(counts), _p(rules), _p(colors), _p(lines), _p(image_grad),
                               _p(line_grad), _p(color_grad))
    return line_grad, color_grad


def _masked_backward(width, height, clear, draws, colors, lines, image_grad, masks, color_maps=False):
    """With clip masks: the composite's adjoint in numpy (_masked_render), then the
    line pass of each draw (or_diff_line_grad_f64) on its dL/dcoverage."""
    lines = np.ascontiguousarray(lines, np.float64).reshape(-1, 4)
    cov_grad, (color_grad, maps) = _masked_render(width, height, clear, draws, colors, lines, masks, image_grad)
    line_grad = np.zeros_like(lines)
    for d, (start, count, rule) in enumerate(draws):
        gmap = np.ascontiguousarray(cov_grad[d][0], np.float64)
        if count and gmap.size:
            lib().or_diff_line_grad_f64(_p(lines), C.c_uint32(start), C.c_uint32(count), C.c_uint32(width),
                                        C.c_uint32(height), _p(gmap), _p(line_grad))
    return line_grad, (maps if color_maps else color_grad)


def sdf_draw_coverage(arrays, draw_index):
    """FP32 SDF coverage of one clip draw at every pixel centre (H, W)."""
    draws = np.ascontiguousarray(arrays.draws).view(np.uint32).reshape(-1)
    paint = np.ascontiguousarray(arrays.paint, np.uint32)
    out = np.zeros((arrays.height, arrays.width), np.float32)
    lib().or_sdf_draw_coverage(_p(draws), C.c_uint32(len(arrays.draws)), C.c_uint32(draw_index), _p(paint),
                               C.c_uint32(paint.size), C.c_uint32(arrays.paint_sdf_shadow_base),
                               C.c_uint32(arrays.width), C.c_uint32(arrays.height), _p(out))
    return out
