"""The GPU front end (tk_frame_set_polygons): new vertices for a polygon scene,
recorded once on the host, rebuild on the device exactly what recording the
same polygons and encoding produces (src/canvas/paths.rs:65-156,
src/shared/gpu_plan/path_plans.rs:156-228,463-504,
src/shared/gpu_plan/tile_bins/build.rs:4-83, transient.rs:7-146): every scene
array word for word against a frame made from the host recording of the new
vertices, including early occlusion by opaque SDF rects, path and SDF clip
layers, degenerate edges, growth and shrinkage of every buffer, and batches of
scenes; the renders then equal the host-recorded frame's bit for bit."""
import numpy as np
import pytest

import tileink_amd as tk
from tileink_amd import _abi

pytestmark = pytest.mark.gpu

SCENE_BUFFERS = [
    ("lines", _abi.BUF_LINES, 6, "lines"), ("paths", _abi.BUF_PATHS, 19, "paths"),
    ("draws", _abi.BUF_DRAWS, 31, "draws"), ("tile_draws", _abi.BUF_TILE_DRAWS, 2, "tiles"),
    ("tile_draw_indices", _abi.BUF_TILE_DRAW_INDICES, 1, "tile_draw_refs"),
    ("scan_chunks", _abi.BUF_SCAN_CHUNKS, 4, "scan_chunk_count"),
    ("scan_chunk_ranges", _abi.BUF_SCAN_CHUNK_RANGES, 2, "paths"),
    ("cumsum_offsets", _abi.BUF_CUMSUM_CHUNK_OFFSETS, 1, "cumsum_chunk_count"),
    ("cumsum_lens", _abi.BUF_CUMSUM_CHUNK_LENS, 1, "cumsum_chunk_count"),
    ("cumsum_row_starts", _abi.BUF_CUMSUM_ROW_STARTS, 1, "cumsum_row_count"),
    ("cumsum_row_ends", _abi.BUF_CUMSUM_ROW_ENDS, 1, "cumsum_row_count"),
]
STAT_KEYS = ("paths", "lines", "draws", "batches", "tiles", "backdrop_len", "segment_capacity", "tile_draw_refs",
             "scan_chunk_count", "cumsum_chunk_count", "cumsum_row_count", "ptcl_bound")


def blob(rng, cx, cy, r, n):
    a = np.sort(rng.uniform(0, 2 * np.pi, n))
    rr = r * rng.uniform(0.4, 1.0, n)
    return np.stack([cx + rr * np.cos(a), cy + rr * np.sin(a)], 1)


class PolygonScene:
    """A recording schedule over polygons: ('poly', k, rgba, rule), ('rect', box,
    radius, rgba), ('clip_sdf', box, radius), ('clip_path', k, rule), ('pop',).
    Path k takes polygon k of whichever vertex set the scene is recorded with."""

    def __init__(self, seed, width, height, count=40, occluders=3, clips=True):
        rng = np.random.default_rng(seed)
        self.width, self.height = width, height
        self.ops, polys = [], []

        def poly():
            n = int(rng.integers(3, 40))
            p = blob(rng, rng.uniform(0, width), rng.uniform(0, height), rng.lognormal(np.log(40), 0.8), n)
            if n > 4 and rng.uniform() < 0.3:  # a repeated vertex: a degenerate edge
                p[2] = p[1]
            polys.append(p.astype(np.float32))
            return len(polys) - 1

        def color(opaque=False):
            return tuple(rng.uniform(0, 1, 3)) + ((1.0,) if opaque else (float(rng.uniform(0.3, 1)),))

        opened = 0
        for i in range(count):
            if clips and i == count // 4:
                self.ops.append(("clip_sdf", (3.5, 5.25, width - 7.0, height - 2.5), 9.0))
                opened += 1
            if clips and i == count // 2:
                k = poly()
                polys[k] = blob(rng, width / 2, height / 2, 0.45 * min(width, height), 24).astype(np.float32)
                self.ops.append(("clip_path", k, int(rng.integers(0, 2))))
                opened += 1
            self.ops.append(("poly", poly(), color(), int(rng.integers(0, 2))))
            if occluders and i % max(1, count // occluders) == count // (2 * max(1, occluders)):
                x0, y0 = rng.uniform(0, width * 0.6), rng.uniform(0, height * 0.6)
                self.ops.append(("rect", (float(np.floor(x0)), float(np.floor(y0)), float(np.floor(x0)) + 96.0,
                                          float(np.floor(y0)) + 80.0), 0.0, color(opaque=True)))
        self.ops.extend([("pop",)] * opened)
        self.polygons = polys

    def record(self, polygons):
        canvas = tk.Canvas(self.width, self.height)
        for op in self.ops:
            if op[0] == "poly":
                _, k, rgba, rule = op
                p = np.asarray(polygons[k], np.float64)
                canvas.push_polygons(p, [len(p)], np.array([rgba], np.float32), [rule])
            elif op[0] == "rect":
                canvas.push_rect(*op[1], op[2], tk.Color(*op[3]))
            elif op[0] == "clip_sdf":
                canvas.push_clip_sdf_rect_layer(*op[1], op[2])
            elif op[0] == "clip_path":
                p = np.asarray(polygons[op[1]], np.float64)
                path = tk.BezPath().move_to(*p[0])
                for q in p[1:]:
                    path.line_to(*q)
                canvas.push_clip_layer(path.close_path(), rule=tk.FillRule(op[2]))
            else:
                canvas.pop_layer()
        return canvas

    def path_order(self):
        """Polygon index of each recorded path, in path order."""
        return [op[1] for op in self.ops if op[0] in ("poly", "clip_path")]

    def points(self, polygons):
        order = self.path_order()
        pts = np.concatenate([polygons[k] for k in order]).astype(np.float32)
        starts = np.concatenate([[0], np.cumsum([len(polygons[k]) for k in order])]).astype(np.uint32)
        return pts, starts

    def moved(self, seed, scale=1.0, sigma=3.0):
        rng = np.random.default_rng(seed)
        out = []
        for p in self.polygons:
            c = p.mean(0)
            q = c + (p - c) * scale + rng.normal(0, sigma, p.shape)
            out.append(q.astype(np.float32))
        return out


def scene_arrays(frame):
    st = frame.stats()
    out = {k: st[k] for k in STAT_KEYS}
    for name, which, words, count_key in SCENE_BUFFERS:
        out["buf_" + name] = frame.buffer(which, np.uint32, words * int(st[count_key]))
    return out


def assert_same_scene(got, want):
    for k in STAT_KEYS:
        assert got[k] == want[k], (k, got[k], want[k])
    for name, *_ in SCENE_BUFFERS:
        a, b = got["buf_" + name], want["buf_" + name]
        assert a.shape == b.shape and np.array_equal(a, b), (name, int(np.count_nonzero(a != b)) if a.shape == b.shape
                                                            else (a.shape, b.shape))


def check_move(ctx, sc, frame, polygons):
    pts, starts = sc.points(polygons)
    lv = frame.set_polygons(pts, starts, line_vertex=True)
    ref = tk.Frame(ctx, sc.record(polygons).encode(0))
    assert_same_scene(scene_arrays(frame), scene_arrays(ref))
    # each line runs from its first vertex to the next vertex of its polygon
    lines = frame.buffer(_abi.BUF_LINES, np.float32, 6 * frame.stats()["lines"]).reshape(-1, 6)
    path_of = np.repeat(np.arange(len(starts) - 1), np.diff(starts))
    nxt = np.where(np.arange(len(pts)) + 1 == starts[path_of + 1], starts[path_of], np.arange(len(pts)) + 1)
    assert np.array_equal(lines[:, 2:4], pts[lv]) and np.array_equal(lines[:, 4:6], pts[nxt[lv]])
    frame.render()
    ref.render()
    assert frame.stats()["overflow"] == 0
    assert np.array_equal(frame.image(), ref.image())


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_set_polygons_rebuilds_the_host_encoding(ctx, seed):
    sc = PolygonScene(seed, 301 + 37 * seed, 203 + 11 * seed, count=60, occluders=4)
    frame = tk.Frame(ctx, sc.record(sc.polygons).encode(0))
    # same geometry, moved, grown past every buffer, then shrunk
    for polygons in (sc.polygons, sc.moved(seed), sc.moved(seed + 10, scale=2.5, sigma=6.0),
                     sc.moved(seed + 20, scale=0.3)):
        check_move(ctx, sc, frame, polygons)


def test_set_polygons_with_wide_canvas_and_no_clips(ctx):
    # more than 256 tiles per row (the row kernel's lanes loop over x)
    sc = PolygonScene(9, 4700, 90, count=200, occluders=6, clips=False)
    frame = tk.Frame(ctx, sc.record(sc.polygons).encode(0))
    check_move(ctx, sc, frame, sc.moved(4, scale=1.3))


def test_set_polygons_on_a_batch(ctx):
    scenes = [PolygonScene(20 + i, w, h, count=n) for i, (w, h, n) in
              enumerate([(128, 96, 12), (400, 250, 70), (64, 300, 20), (333, 333, 45)])]
    batch = tk.Batch(ctx, [s.record(s.polygons).encode(0) for s in scenes])
    moved = [s.moved(50 + i, scale=1.5) for i, s in enumerate(scenes)]
    parts = [s.points(m) for s, m in zip(scenes, moved)]
    pts = np.concatenate([p for p, _ in parts])
    offs = np.cumsum([0] + [len(p) for p, _ in parts[:-1]])
    starts = np.concatenate([st[:-1] + o for (_, st), o in zip(parts, offs)] + [[len(pts)]]).astype(np.uint32)
    batch.frame.set_polygons(pts, starts)
    ref = tk.Batch(ctx, [s.record(m).encode(0) for s, m in zip(scenes, moved)])
    assert_same_scene(scene_arrays(batch.frame), scene_arrays(ref.frame))
    batch.render()
    ref.render()
    for i in range(len(scenes)):
        assert np.array_equal(batch.image(i), ref.image(i)), i


def test_set_polygons_refuses_non_finite_vertices(ctx):
    sc = PolygonScene(5, 200, 150, count=10)
    frame = tk.Frame(ctx, sc.record(sc.polygons).encode(0))
    pts, starts = sc.points(sc.polygons)
    pts[3, 0] = np.nan
    with pytest.raises(tk.TileinkError) as err:
        frame.set_polygons(pts, starts)
    assert err.value.status == _abi.TK_ERR_INVALID
    check_move(ctx, sc, frame, sc.polygons)  # valid vertices restore the frame


class CubicScene:
    """Closed contours of 4-16 cubics (the clipped-paths generator's fills), with
    opaque SDF rects, an SDF clip and a path clip made of cubics."""

    def __init__(self, seed, width, height, count=80):
        rng = np.random.default_rng(seed)
        self.width, self.height = width, height
        self.contours, self.ops = [], []
        for i in range(count):
            if i == count // 3:
                self.ops.append(("clip_sdf", (5.5, 3.25, width - 9.0, height - 4.5), 11.0))
            if i == count // 2:
                self.ops.append(("clip_path", self._contour(rng, width / 2, height / 2, 0.4 * min(width, height), 6),
                                 int(rng.integers(0, 2))))
            c = self._contour(rng, rng.uniform(0, width), rng.uniform(0, height),
                              float(rng.lognormal(np.log(48.0), 1.0)), int(rng.integers(4, 17)))
            color = tuple(rng.uniform(0, 1, 3)) + (float(rng.uniform(0.3, 1)),)
            self.ops.append(("path", c, color, int(rng.integers(0, 2))))
            if i % 20 == 7:
                x0, y0 = float(np.floor(rng.uniform(0, width * 0.7))), float(np.floor(rng.uniform(0, height * 0.7)))
                self.ops.append(("rect", (x0, y0, x0 + 80.0, y0 + 64.0), 0.0, tuple(rng.uniform(0, 1, 3)) + (1.0,)))
        self.ops.extend([("pop",), ("pop",)])

    def _contour(self, rng, cx, cy, extent, n):
        pts = rng.uniform(-extent / 2, extent / 2, size=(3 * n + 1, 2)) + (cx, cy)
        pts[-1] = pts[0]  # the last cubic returns to the start
        k = len(self.contours)
        self.contours.append(pts)
        return k

    def record(self, contours):
        canvas = tk.Canvas(self.width, self.height)

        def path(pts):
            p = tk.BezPath().move_to(*pts[0])
            for i in range(1, len(pts), 3):
                p.curve_to(*pts[i], *pts[i + 1], *pts[i + 2])
            return p.close_path()

        for op in self.ops:
            if op[0] == "path":
                canvas.push_path(path(contours[op[1]]), tk.Color(*op[2]), rule=tk.FillRule(op[3]), tolerance=0.25)
            elif op[0] == "rect":
                canvas.push_rect(*op[1], op[2], tk.Color(*op[3]))
            elif op[0] == "clip_sdf":
                canvas.push_clip_sdf_rect_layer(*op[1], op[2])
            elif op[0] == "clip_path":
                canvas.push_clip_layer(path(contours[op[1]]), rule=tk.FillRule(op[2]), tolerance=0.25)
            else:
                canvas.pop_layer()
        return canvas

    def ctrl(self, contours):
        order = [op[1] for op in self.ops if op[0] in ("path", "clip_path")]
        pts = np.concatenate([contours[k] for k in order])
        starts = np.concatenate([[0], np.cumsum([len(contours[k]) for k in order])]).astype(np.uint32)
        return pts, starts

    def moved(self, seed, sigma=5.0, scale=1.0):
        rng = np.random.default_rng(seed)
        out = []
        for c in self.contours:
            m = c.mean(0)
            q = m + (c - m) * scale + rng.normal(0, sigma, c.shape)
            q[-1] = q[0]
            out.append(q)
        return out


@pytest.mark.parametrize("seed", [7, 8])
def test_set_cubic_paths_flattens_on_the_device_like_the_recorder(ctx, seed):
    # GPU flattening (flatten.h on the device) + the front end, against the host
    # recorder's flattening and encoding of the same control points: every scene
    # array word for word, then the images.
    sc = CubicScene(seed, 640, 480)
    frame = tk.Frame(ctx, sc.record(sc.contours).encode(0))
    for contours in (sc.contours, sc.moved(seed), sc.moved(seed + 1, sigma=2.0, scale=2.0)):
        ctrl, starts = sc.ctrl(contours)
        frame.set_cubic_paths(ctrl, starts, tolerance=0.25)
        ref = tk.Frame(ctx, sc.record(contours).encode(0))
        assert_same_scene(scene_arrays(frame), scene_arrays(ref))
        frame.render()
        ref.render()
        assert frame.stats()["overflow"] == 0
        assert np.array_equal(frame.image(), ref.image())


def test_set_cubic_paths_refuses_malformed_contours(ctx):
    sc = CubicScene(3, 200, 160, count=10)
    frame = tk.Frame(ctx, sc.record(sc.contours).encode(0))
    ctrl, starts = sc.ctrl(sc.contours)
    bad = starts.copy()
    bad[1] += 1  # path 0 would hold 3k + 2 points
    with pytest.raises(tk.TileinkError) as err:
        frame.set_cubic_paths(ctrl, bad)
    assert err.value.status == _abi.TK_ERR_INVALID
    ctrl[5, 1] = np.inf
    with pytest.raises(tk.TileinkError) as err:
        frame.set_cubic_paths(ctrl, starts)
    assert err.value.status == _abi.TK_ERR_INVALID
