"""Host recording and planning (CPU): flatten, stroke, records, batches, bins, chunk plans.

Ports the reference's own recording tests (src/shared/path_flatten.rs:84-174,
src/canvas/tests/paths.rs:29-139) and checks the geometry front end against
the mathematics it must satisfy (chord deviation, curve membership).
"""
import math

import numpy as np
import pytest

import tileink_amd as tk
from tileink_amd import scenes


def lines_of(canvas):
    a = canvas.arrays()
    return a.lines, a


def test_line_after_move_to_is_not_skipped():
    c = tk.Canvas(64, 64)
    c.push_path(tk.BezPath().move_to(1, 2).line_to(5, 2), tk.Color(1, 0, 0), tolerance=0.1)
    lines, _ = lines_of(c)
    assert len(lines) == 2
    assert tuple(lines[0]["p0"]) == (1.0, 2.0) and tuple(lines[0]["p1"]) == (5.0, 2.0)
    assert tuple(lines[1]["p0"]) == (5.0, 2.0) and tuple(lines[1]["p1"]) == (1.0, 2.0)


def test_final_open_contour_is_closed():
    c = tk.Canvas(64, 64)
    c.push_path(tk.BezPath().move_to(0, 0).line_to(10, 0).line_to(10, 10), tk.Color(1, 0, 0), tolerance=0.1)
    lines, _ = lines_of(c)
    assert len(lines) == 3
    assert tuple(lines[2]["p0"]) == (10.0, 10.0) and tuple(lines[2]["p1"]) == (0.0, 0.0)


def test_closed_rect_keeps_all_edges():
    c = tk.Canvas(64, 64)
    c.push_path(tk.BezPath().move_to(0, 0).line_to(10, 0).line_to(10, 10).line_to(0, 10).close_path(),
                tk.Color(1, 0, 0), tolerance=0.1)
    assert len(lines_of(c)[0]) == 4


def test_duplicate_points_and_repeated_close_preserve_only_real_edges():
    c = tk.Canvas(64, 64)
    for _ in range(7):  # make the path id 7 as in the reference case
        c.push_path(tk.BezPath().move_to(50, 50).line_to(51, 50).line_to(51, 51).close_path(), tk.Color(0, 0, 0))
    p = tk.BezPath().move_to(0, 0).line_to(0, 0).line_to(10, 0).line_to(10, 0).close_path().close_path()
    c.push_path(p, tk.Color(1, 0, 0), tolerance=0.1)
    lines, _ = lines_of(c)
    mine = lines[lines["path_id"] == 7]
    assert len(mine) == 2
    assert (tuple(mine[0]["p0"]), tuple(mine[0]["p1"])) == ((0.0, 0.0), (10.0, 0.0))
    assert (tuple(mine[1]["p0"]), tuple(mine[1]["p1"])) == ((10.0, 0.0), (0.0, 0.0))


def _circle_path(cx, cy, r):
    k = 0.5522847498307936 * r
    p = tk.BezPath().move_to(cx + r, cy)
    p.curve_to(cx + r, cy + k, cx + k, cy + r, cx, cy + r)
    p.curve_to(cx - k, cy + r, cx - r, cy + k, cx - r, cy)
    p.curve_to(cx - r, cy - k, cx - k, cy - r, cx, cy - r)
    p.curve_to(cx + k, cy - r, cx + r, cy - k, cx + r, cy)
    return p.close_path()


def test_circle_flattens_to_many_small_lines():
    c = tk.Canvas(64, 64)
    c.push_path(_circle_path(20, 20, 10), tk.Color(1, 0, 0), tolerance=0.1)
    lines, _ = lines_of(c)
    assert len(lines) > 16
    pts = np.concatenate([lines["p0"], lines["p1"]])
    assert pts.min() >= 9.0 and pts.max() <= 31.0


def _bezier(pts, t):
    t = t[:, None]
    return ((1 - t) ** 3) * pts[0] + 3 * (1 - t) ** 2 * t * pts[1] + 3 * (1 - t) * t ** 2 * pts[2] + t ** 3 * pts[3]


def _dist_to_polyline(samples, poly):
    a0, a1 = poly[:-1], poly[1:]
    best = np.full(len(samples), np.inf)
    for p, q in zip(a0, a1):
        d = q - p
        t = np.clip(((samples - p) @ d) / max(d @ d, 1e-300), 0, 1)
        best = np.minimum(best, np.hypot(*(samples - (p + t[:, None] * d)).T))
    return best


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("tol", [0.25, 0.05])
def test_cubic_flattening_within_tolerance_and_on_curve(seed, tol):
    """The curve stays within `tol` of the chords, and chord vertices lie on the
    quadratic approximations, i.e. within a tenth of `tol` of the cubic."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 500, (4, 2))
    c = tk.Canvas(512, 512)
    c.push_path(tk.BezPath().move_to(*pts[0]).curve_to(*pts[1], *pts[2], *pts[3]), tk.Color(1, 0, 0), tolerance=tol)
    lines, _ = lines_of(c)
    poly = np.concatenate([lines["p0"][:-1], lines["p1"][-2:-1]]).astype(np.float64)
    assert np.allclose(poly[0], pts[0], atol=1e-4) and np.allclose(poly[-1], pts[3], atol=1e-4)
    dense = _bezier(pts, np.linspace(0, 1, 4001))
    assert _dist_to_polyline(dense, poly).max() <= tol + 1e-3
    # every vertex is within 0.1 tol of the curve (plus f32 storage rounding)
    assert _dist_to_polyline(poly, dense).max() <= 0.1 * tol + 2e-3


def test_stroke_bounds_cover_shape_plus_half_width():
    # push_stroke_expands_shape_to_fill_path (src/canvas/tests/paths.rs:29-49)
    c = tk.Canvas(64, 64)
    rect = tk.BezPath().move_to(10, 10).line_to(20, 10).line_to(20, 20).line_to(10, 20).close_path()
    c.push_stroke(rect, 4.0, tk.Color(1, 0, 0), tolerance=0.1)
    a = c.arrays()
    x0, y0, x1, y1 = [int(v) for v in a.draws[0, 10:14].view(np.int32)]
    assert x0 <= 8 and y0 <= 8 and x1 >= 22 and y1 >= 22
    assert a.draws[0, 9] == tk.FILL_NONZERO


def test_stroke_outline_stays_within_half_width():
    """Every outline vertex is at most w/2 + tol from the centreline, and the outline
    reaches out to w/2 around it."""
    c = tk.Canvas(256, 256)
    pts = np.array([[20.0, 30.0], [200.0, 60.0], [40.0, 220.0], [230.0, 200.0]])
    c.push_stroke(tk.BezPath().move_to(*pts[0]).curve_to(*pts[1], *pts[2], *pts[3]), 6.0, tk.Color(0, 0, 0),
                  tolerance=0.1)
    lines, _ = lines_of(c)
    outline = np.concatenate([lines["p0"], lines["p1"]]).astype(np.float64)
    center = _bezier(pts, np.linspace(0, 1, 6001))
    d = _dist_to_polyline(outline, center)
    assert d.max() <= 3.0 + 0.1 + 1e-3
    assert np.percentile(d, 50) >= 3.0 - 0.1 - 1e-3


def test_translated_path_bounds_and_scale():
    # push_path_flattens_transformed_geometry / new_scales_path_geometry_before_flattening
    rect = tk.BezPath().move_to(0, 0).line_to(10, 0).line_to(10, 10).line_to(0, 10).close_path()
    c = tk.Canvas(64, 64)
    c.push_path(rect, tk.Color(1, 0, 0), transform=(1, 0, 0, 1, 8, 4))
    a = c.arrays()
    assert list(a.draws[0, 10:14].view(np.int32)) == [8, 4, 18, 14]
    pts = np.concatenate([a.lines["p0"], a.lines["p1"]])
    assert pts[:, 0].min() >= 8 and pts[:, 0].max() <= 18 and pts[:, 1].min() >= 4 and pts[:, 1].max() <= 14
    c2 = tk.Canvas(64, 64, 2.0)
    c2.push_path(rect, tk.Color(1, 0, 0), transform=(1, 0, 0, 1, 8, 4))
    assert list(c2.arrays().draws[0, 10:14].view(np.int32)) == [16, 8, 36, 28]


def test_segment_capacity_is_sum_of_scanned_tile_counts():
    # push_path_reserves_segment_capacity_from_scan_tile_count (src/canvas/tests/paths.rs:104-139)
    c = tk.Canvas(64, 64)
    c.push_path(tk.BezPath().move_to(8, 8).line_to(9, 12), tk.Color(1, 0, 0))
    a = c.arrays()
    rec = a.paths[0]
    bbox = rec[6:10]
    expected = sum(tk.line_scanned_tile_count(tuple(l["p0"]), tuple(l["p1"]), bbox) for l in a.lines)
    assert rec[11] == expected and a.segment_capacity == expected
    assert 0 < expected < 20


def test_clip_layers_compile_into_batches_with_stack_snapshots():
    c = tk.Canvas(128, 128)
    red = tk.Color(1, 0, 0)
    c.push_rect(0, 0, 10, 10, 0, red)           # batch 0: no stack
    c.push_clip_sdf_rect_layer(5, 5, 100, 100, 4)
    c.push_rect(0, 0, 20, 20, 0, red)           # batch 1: [clip A]
    c.push_clip_sdf_rect_layer(10, 10, 90, 90)
    c.pop_layer()                               # empty child list: no batch
    c.push_rect(0, 0, 30, 30, 0, red)           # batch 1 again (same stack) -> coalesced
    c.pop_layer()
    c.push_rect(0, 0, 40, 40, 0, red)           # batch 2: no stack
    a = c.arrays()
    assert a.batch_count == 3
    assert a.batches.tolist() == [[0, 0], [0, 1], [1, 1]]
    assert a.layers[:, 0].tolist() == [0]       # LAYER_CLIP
    ids = a.draw_batch_ids
    assert ids.tolist() == [0, tk.INVALID, 1, tk.INVALID, 1, 2]
    with pytest.raises(tk.TileinkError):
        c.pop_layer()


def test_pixel_aligned_path_clip_becomes_sdf_clip():
    c = tk.Canvas(64, 64)
    c.push_clip_layer(tk.BezPath().move_to(8, 8).line_to(40, 8).line_to(40, 40).line_to(8, 40).close_path())
    c.push_rect(0, 0, 64, 64, 0, tk.Color(0, 0, 1))
    c.pop_layer()
    a = c.arrays()
    assert a.path_count == 0  # no path record: rasterised as an SDF rectangle
    clip = a.layers[0, 1]
    assert a.draws[clip, 8] == 1  # DRAW_CLIP
    assert a.draws[clip, 2] != tk.INVALID  # SDF record


def test_tile_bins_are_painter_ordered_per_tile():
    c = scenes.mixed_scene_640x360()
    a = c.arrays()
    order = np.flatnonzero(a.draw_batch_ids != tk.INVALID)
    rank = np.empty(a.draw_count, np.int64)
    rank[order] = np.arange(len(order))
    for t in range(0, a.tiles_width * a.tiles_height, 37):
        start, n = a.tile_draws[t]
        if n == 0:
            assert start == tk.INVALID
            continue
        draws = a.tile_draw_indices[start:start + n]
        assert np.all(np.diff(rank[draws]) > 0)


def test_scan_and_cumsum_chunk_plans_partition_each_path():
    a = scenes.clipped_paths_small(16).arrays()
    for p in range(a.path_count):
        rec = a.paths[p]
        s, e = a.scan_chunk_ranges[p]
        chunks = a.scan_chunks[s:e]
        assert chunks[:, 3].sum() == rec[5]
        assert np.all(chunks[:, 3] <= 256) and np.all(chunks[:, 0] == p)
        if len(chunks):
            assert chunks[0, 1] == rec[4]
            assert np.all(np.diff(chunks[:, 1]) == chunks[:-1, 3])
    assert np.all(a.cumsum_chunk_lens <= 256)
    assert a.cumsum_chunk_lens.sum() == a.backdrop_len


def test_invalid_arguments_are_rejected():
    c = tk.Canvas(16, 16)
    with pytest.raises(tk.TileinkError):
        c.push_path(tk.BezPath().move_to(0, 0).line_to(1, 1), tk.Color(1, 0, 0), rule=5)
    with pytest.raises(tk.TileinkError):
        c.push_path(tk.BezPath().move_to(0, 0).line_to(float("nan"), 1), tk.Color(1, 0, 0))
    with pytest.raises(tk.TileinkError):
        tk.Canvas(0, 10)


def _covered_area(canvas):
    import oracle
    image, _ = oracle.render_frame(canvas.arrays(0))
    return (image.view(np.uint8).reshape(image.shape + (4,))[..., 3].astype(np.float64) / 255.0).sum()


def test_stroke_joins_and_caps_add_their_closed_form_areas():
    # A V of two 72.1 px segments, width 10 (h = 5), turning by 112.6 degrees at the joint.
    def area(join, cap):
        c = tk.Canvas(120, 100)
        c.push_stroke(tk.BezPath().move_to(20, 20).line_to(60, 80).line_to(100, 20), 10.0, tk.Color(0, 0, 0, 1),
                      join=join, start_cap=cap, end_cap=cap)
        return _covered_area(c)
    h = 5.0
    turn = math.pi - math.acos(((-40) * 40 + (-60) * (-60)) / (40 ** 2 + 60 ** 2))
    bevel = area(tk.Join.Bevel, tk.Cap.Butt)
    # square caps add an h x 2h rectangle at each end
    assert abs(area(tk.Join.Bevel, tk.Cap.Square) - bevel - 2 * 2 * h * h) < 0.5
    # the miter triangle beyond the bevel: h^2 (tan(t/2) - sin(t/2) cos(t/2))
    miter_extra = h * h * (math.tan(turn / 2) - math.sin(turn / 2) * math.cos(turn / 2))
    assert abs(area(tk.Join.Miter, tk.Cap.Butt) - bevel - miter_extra) < 0.5
    # the round join adds the circular segment h^2 (t - sin t) / 2 (its polygon slightly less)
    round_extra = h * h * (turn - math.sin(turn)) / 2
    assert 0.9 * round_extra < area(tk.Join.Round, tk.Cap.Butt) - bevel < round_extra + 0.3
    # a miter limit below 1 / cos(t/2) turns the miter into a bevel
    c = tk.Canvas(120, 100)
    c.push_stroke(tk.BezPath().move_to(20, 20).line_to(60, 80).line_to(100, 20), 10.0, tk.Color(0, 0, 0, 1),
                  join=tk.Join.Miter, start_cap=tk.Cap.Butt, end_cap=tk.Cap.Butt, miter_limit=1.5)
    assert abs(_covered_area(c) - bevel) < 0.05


def test_default_stroke_is_round_joins_and_caps():
    path = tk.BezPath().move_to(20, 20).line_to(60, 80).line_to(100, 20)
    a, b = tk.Canvas(120, 100), tk.Canvas(120, 100)
    a.push_stroke(path, 6.0, tk.Color(0, 0, 0, 1))
    b.push_stroke(path, 6.0, tk.Color(0, 0, 0, 1), join=tk.Join.Round, start_cap=tk.Cap.Round, end_cap=tk.Cap.Round)
    la, lb = a.arrays(0).lines, b.arrays(0).lines
    assert np.array_equal(la, lb)


def test_invalid_stroke_styles_are_rejected():
    c = tk.Canvas(64, 64)
    path = tk.BezPath().move_to(10, 10).line_to(50, 50)
    with pytest.raises(tk.TileinkError):
        c.push_stroke(path, 4.0, tk.Color(0, 0, 0, 1), join=7)
    with pytest.raises(tk.TileinkError):
        c.push_stroke(path, 4.0, tk.Color(0, 0, 0, 1), join=tk.Join.Miter, miter_limit=0.5)


def test_bulk_polygons_record_what_single_polygons_record():
    rng = np.random.default_rng(21)
    polys = [rng.uniform(5, 120, (int(rng.integers(3, 12)), 2)) for _ in range(9)]
    colors = rng.uniform(0, 1, (9, 4)).astype(np.float32)
    rules = rng.integers(0, 2, 9)
    a, b = tk.Canvas(128, 128), tk.Canvas(128, 128)
    for p, c, r in zip(polys, colors, rules):
        a.push_polygon(p, tk.Color(*c), tk.FillRule(int(r)))
    ids = b.push_polygons(np.concatenate(polys), [len(p) for p in polys], colors, rules)
    assert list(ids) == list(range(9))
    sa, sb = a.arrays(0), b.arrays(0)
    for name in ("lines", "paths", "draws", "paint", "tile_draw_indices"):
        assert np.array_equal(np.asarray(getattr(sa, name)), np.asarray(getattr(sb, name))), name
    with pytest.raises(tk.TileinkError):
        b.push_polygons(np.zeros((4, 2)), [3], colors[:1], rules[:1])  # counts disagree with the points
