"""The oracle pinned against the reference's known-answer tests and closed forms (CPU).

KATs restated from the reference's GPU tests:
  scan prefix      src/native/runtime/tests/scan_cases.rs:8-157, metal/scan_tests.rs:3-22
  scan count       src/native/runtime/tests/scan_cases.rs:159-214, metal/scan_tests.rs:24-95
  coarse prefix    src/native/runtime/tests/coarse_cases.rs:4-75, metal/coarse_tests.rs:5-19
  coarse count     src/native/runtime/tests/coarse_count/scene.rs:7-106, coarse_count_gpu.rs:11-30
  pixel rounding   src/native/runtime/tests/pixel_rounding_gpu.rs:22-27
Closed forms and brute force pin what those leave open (tile crossings, exact
pixel areas, winding numbers).
"""
import ctypes as C
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import scenes

GUARD = 0x5555_5555


def lcg(seed):
    return (seed * 1_664_525 + 1_013_904_223) & 0xFFFFFFFF


# ---------------------------------------------------------------------------
# scan prefix
# ---------------------------------------------------------------------------

def prefix_case(lengths, seed):
    """Inputs and expected outputs of scan_cases::prefix_case."""
    chunks, counts, cursor, rng = [], [0x9876_5432] * 7, 7, seed
    for n in lengths:
        chunks.append((0, cursor, 19, n))
        for _ in range(n):
            rng = lcg(rng)
            counts.append(rng)
        counts += [0x9876_5432] * 3
        cursor = len(counts)
    ranges = [GUARD] * (2 * len(counts))
    cursors = [GUARD] * len(counts)
    totals, offsets, nxt = [], [], 19
    for (_, off, _, n) in chunks:
        offsets.append(nxt)
        total = 0
        for i in range(off, off + n):
            ranges[2 * i] = (nxt + total) & 0xFFFFFFFF
            cursors[i] = (nxt + total) & 0xFFFFFFFF
            total = (total + counts[i]) & 0xFFFFFFFF
            ranges[2 * i + 1] = (nxt + total) & 0xFFFFFFFF
        totals.append(total)
        nxt = (nxt + total) & 0xFFFFFFFF
    inputs = {
        "config": [len(counts), len(counts), 2, len(lengths), 0, 0, 0, 0, 0, 0, 0],
        "chunks": np.array(chunks, np.uint32).reshape(-1, 4),
        "paths": np.array([[0, 0, 0, 0, 7, len(counts) - 7, 0, 0, 0, 0, 19, 0, 0] + [0] * 6,
                           [1] + [0] * 18], np.uint32),
        "chunk_ranges": np.array([[0, len(lengths)], [len(lengths), len(lengths)]], np.uint32),
        "counts": np.array(counts, np.uint32),
    }
    expected = {
        "ranges": np.array(ranges, np.uint32), "cursors": np.array(cursors, np.uint32),
        "totals": np.array(totals, np.uint32), "offsets": np.array(offsets, np.uint32),
        "bumps": np.array([(nxt - 19) & 0xFFFFFFFF, 0], np.uint32), "counts": np.array(counts, np.uint32),
    }
    return inputs, expected


PREFIX_LENGTHS = [[0], [1], [255], [256], [256, 1], [256, 0, 17, 1, 255]]
PREFIX_SEEDS = [0, 19, 0xFFFFFFFF]


@pytest.mark.parametrize("lengths", PREFIX_LENGTHS)
@pytest.mark.parametrize("seed", PREFIX_SEEDS)
def test_oracle_scan_prefix_kat(lengths, seed):
    inp, exp = prefix_case(lengths, seed)
    cfg = oracle.ScanConfig(*inp["config"])
    n = len(inp["counts"])
    ranges = np.full(2 * n, GUARD, np.uint32)
    cursors = np.full(n, GUARD, np.uint32)
    totals = np.zeros(max(len(lengths), 1), np.uint32)
    offsets = np.zeros(max(len(lengths), 1), np.uint32)
    bumps = np.full(2, GUARD, np.uint32)
    counts = inp["counts"].copy()
    oracle.scan_prefix(cfg, inp["chunks"], inp["chunk_ranges"], inp["paths"], counts, ranges, cursors, bumps,
                       totals, offsets)
    assert np.array_equal(ranges, exp["ranges"])
    assert np.array_equal(cursors, exp["cursors"])
    assert np.array_equal(totals[:len(lengths)], exp["totals"])
    assert np.array_equal(offsets[:len(lengths)], exp["offsets"])
    assert np.array_equal(bumps, exp["bumps"])
    assert np.array_equal(counts, exp["counts"])


# ---------------------------------------------------------------------------
# scan count
# ---------------------------------------------------------------------------

def count_segments():
    segs = [((8.0, 0.0), (8.0, 48.0)), ((0.0, 8.0), (48.0, 8.0)), ((-8.0, 0.0), (-8.0, 48.0)),
            ((0.0, 0.0), (48.0, 48.0)), ((16.0, 16.0), (16.0, 16.0)), ((0.0, 16.0), (128.0, 16.0)),
            ((-16.0, -16.0), (144.0, 128.0)), ((0.0, 0.000_001), (128.0, 0.000_001))]
    seed = 127
    for _ in range(120):
        pt = []
        for _ in range(2):
            xy = []
            for _ in range(2):
                seed = lcg(seed)
                xy.append(np.float32(seed % 3073) / np.float32(16.0) - np.float32(32.0))
            pt.append(tuple(float(v) for v in xy))
        segs.append((pt[0], pt[1]))
    out = []
    for a, b in segs:
        out += [(a, b), (b, a)]
    return out


def count_case_buffers(p0, p1):
    lines = np.zeros(6, np.uint32)
    lines[2:6] = np.array([p0[0], p0[1], p1[0], p1[1]], np.float32).view(np.uint32)
    path = np.zeros(19, np.uint32)
    path[1], path[4], path[5], path[8], path[9] = 1, 7, 64, 8, 8
    path[13], path[16] = np.float32(1.0).view(np.uint32), np.float32(1.0).view(np.uint32)
    init = np.array([0x1234_5678] * 7 + [0] * 64 + [0x1234_5678] * 7, np.uint32)
    cfg = [0, 64, 1, 0, 1, 0, 0, 0, 0, 0, 0]
    return cfg, lines, path, init


def tiles_crossed_brute_force(p0, p1, bbox):
    """Tiles of the bbox whose open square meets the open segment (exact rationals).
    Returns None when the segment touches a tile corner or runs along / ends on a tile edge,
    where the scan has its own tie-breaking conventions."""
    (x0, y0), (x1, y1) = [(Fraction(float(v[0])) / 16, Fraction(float(v[1])) / 16) for v in (p0, p1)]
    if (x0, y0) == (x1, y1):
        return 0
    for v in (x0, y0, x1, y1):
        if v.denominator == 1:
            return None
    dx, dy = x1 - x0, y1 - y0
    # parameters where the segment meets grid lines
    ts = set()
    for k in range(-4, 14):
        if dx != 0:
            t = (k - x0) / dx
            if 0 < t < 1:
                ts.add(t)
                yv = y0 + t * dy
                if yv.denominator == 1:
                    return None  # passes a corner
        if dy != 0:
            t = (k - y0) / dy
            if 0 < t < 1:
                ts.add(t)
    if dx == 0 or dy == 0:
        # axis-parallel: fine as long as it does not run along a grid line (checked above via endpoints)
        pass
    ts = sorted(ts | {Fraction(0), Fraction(1)})
    tiles = set()
    for a, b in zip(ts, ts[1:]):
        m = (a + b) / 2
        tx, ty = x0 + m * dx, y0 + m * dy
        tiles.add((int(np.floor(tx)), int(np.floor(ty))))
    bx0, by0, bx1, by1 = bbox
    return sum(1 for (tx, ty) in tiles if bx0 <= tx < bx1 and by0 <= ty < by1)


def test_oracle_scan_count_kat_matches_cpu_counts_and_keeps_guards():
    checked_brute = 0
    for p0, p1 in count_segments():
        cfg, lines, path, init = count_case_buffers(p0, p1)
        backdrops = init.copy().view(np.int32)
        counts = init.copy()
        oracle.scan_count(oracle.ScanConfig(*cfg), lines, path, 1, backdrops, counts)
        for buf in (backdrops.view(np.uint32), counts):
            assert np.all(buf[:7] == 0x1234_5678) and np.all(buf[71:] == 0x1234_5678)
        total = int(counts[7:71].sum())
        cpu = oracle.line_scanned_tile_count(p0, p1, (0, 0, 8, 8))
        product = tk.line_scanned_tile_count(p0, p1, (0, 0, 8, 8))
        assert total == cpu == product, (p0, p1)
        brute = tiles_crossed_brute_force(p0, p1, (0, 0, 8, 8))
        if brute is not None:
            assert total == brute, (p0, p1)
            checked_brute += 1
    assert checked_brute >= 150


def test_oracle_backdrops_give_winding_of_segmentless_tiles():
    """After the row prefix, a tile with no segments has one winding number; it must
    equal the path's winding number at the tile centre (f64 ray casting)."""
    canvas = scenes.mixed_scene_640x360()
    a = canvas.arrays()
    _, out = oracle.render_frame(a)
    lines = a.lines
    rng = np.random.default_rng(3)
    for p in rng.choice(a.path_count, 40, replace=False):
        rec = a.paths[p]
        x0, y0, x1, y1 = rec[6:10]
        mine = lines[lines["path_id"] == p]
        q0 = mine["p0"].astype(np.float64)
        q1 = mine["p1"].astype(np.float64)
        for ty in range(y0, y1):
            for tx in range(x0, x1):
                b = rec[4] + (ty - y0) * (x1 - x0) + (tx - x0)
                r0, r1 = out["segment_ranges"][b]
                if r0 != r1:
                    continue
                cx, cy = tx * 16 + 8.0, ty * 16 + 8.0
                # winding: edges crossing the horizontal ray to the left of the centre
                ya, yb = q0[:, 1], q1[:, 1]
                cross = (np.minimum(ya, yb) <= cy) & (np.maximum(ya, yb) > cy)
                t = (cy - ya[cross]) / (yb[cross] - ya[cross])
                xs = q0[cross, 0] + t * (q1[cross, 0] - q0[cross, 0])
                left = xs < cx
                down = (yb[cross] > ya[cross])[left]
                winding = int(np.sum(np.where(down, -1, 1)))
                assert out["backdrops"][b] == winding, (p, tx, ty)


# ---------------------------------------------------------------------------
# coarse prefix (dense tile order)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("count", [0, 1, 255, 256, 257, 65537])
def test_oracle_coarse_prefix_kat(count):
    tiles = count
    chunk_count = (count + 255) // 256
    work = np.full(tiles * 6 + 5 + count + 7, 0x5a5a_5a5a, np.uint32)
    ptcl = np.array([0, 1, 0xFFFFFFFF, 0x8000_0001, 19], np.uint32)
    glyph = np.array([5, 0xFFFFFFFF, 0, 257, 1], np.uint32)
    idx = np.arange(count)
    work[idx * 6] = ptcl[idx % 5]
    work[idx * 6 + 3] = glyph[idx % 5]
    expected = work.copy()
    chunks_expected = np.full((chunk_count + 2) * 4, 0x5a5a_5a5a, np.uint32)
    carry = [0, 0]
    for c in range(chunk_count):
        start = list(carry)
        for t in range(c * 256, min(count, (c + 1) * 256)):
            for kind in range(2):
                expected[t * 6 + kind * 3 + 1] = carry[kind]
                carry[kind] = (carry[kind] + int(work[t * 6 + kind * 3])) & 0xFFFFFFFF
                expected[t * 6 + kind * 3 + 2] = carry[kind]
        chunks_expected[c * 4:c * 4 + 4] = [(carry[0] - start[0]) & 0xFFFFFFFF, start[0],
                                            (carry[1] - start[1]) & 0xFFFFFFFF, start[1]]
    chunks = np.full((chunk_count + 2) * 4, 0x5a5a_5a5a, np.uint32)
    oracle.coarse_prefix(tiles, work, chunks)
    assert np.array_equal(work, expected)
    assert np.array_equal(chunks, chunks_expected)


# ---------------------------------------------------------------------------
# coarse count (count_scene)
# ---------------------------------------------------------------------------

def count_scene(draw_count, relocated, stack):
    pattern = [1, 1, 1, 1, 0, 0, 1, 0]
    draws = np.zeros((8, 31), np.uint32)
    for i in range(8):
        d = draws[i]
        d[0], d[1], d[2], d[4], d[12], d[13] = 0, tk.INVALID, tk.INVALID, tk.INVALID, 16, 16
        d[19] = d[22] = np.float32(1.0).view(np.uint32)
        if i == 0:
            d[2] = 0
        elif i == 1:
            d[4] = 0
        elif i == 3:
            d[8] = 5
        elif i == 4:
            d[8] = 2
        elif i == 6:
            d[8] = 1
        elif i == 7:
            d[0] = tk.INVALID
    prefix = 13 if relocated else 0
    indices = np.full(prefix + draw_count, 0x4545_4545, np.uint32)
    indices[prefix:] = np.arange(draw_count) % 8
    particles = sum(pattern[i % 8] for i in range(draw_count))
    wrapper = 1 if stack == 2 else 0
    if stack == 3:
        particles = 0
    expected = 0 if particles == 0 else particles + 2 * wrapper + 1
    path = np.zeros(19, np.uint32)
    path[8] = path[9] = 1
    layers = np.array([[0 if stack in (1, 3) else (1 if stack == 2 else 99), 2, 0]], np.uint32)
    return {
        "draws": draws, "paths": path, "backdrops": np.array([1 if stack != 3 else 0], np.int32),
        "ranges": np.zeros(2, np.uint32), "layers": layers, "paint": np.zeros(9, np.uint32),
        "batches": np.array([42, 42, 42, 42, 42, 99, 42, 42], np.uint32),
        "tile_draws": np.array([tk.INVALID if draw_count == 0 else prefix, draw_count], np.uint32),
        "indices": indices, "stack": stack, "expected": expected,
    }


@pytest.mark.parametrize("count", [0, 1, 255, 256, 257, 513])
@pytest.mark.parametrize("relocated", [False, True])
@pytest.mark.parametrize("stack", range(4))
def test_oracle_coarse_count_kat(count, relocated, stack):
    sc = count_scene(count, relocated, stack)
    cfg = oracle.CoarseConfig(tile_count=1, tiles_width=1, tiles_height=1, batch_id=42, layer_stack_start=0,
                              layer_stack_end=1 if stack != 0 else 0, ptcl_capacity=3, paint_brush_base=0,
                              draw_count=8, path_count=1, paint_words=9, layer_count=1,
                              tile_draw_index_count=len(sc["indices"]))
    keep = {k: np.ascontiguousarray(v) for k, v in sc.items() if isinstance(v, np.ndarray)}
    inputs = oracle.CoarseInputs(*[keep[k].ctypes.data for k in
                                   ("draws", "paths", "backdrops", "ranges", "layers", "paint", "batches",
                                    "tile_draws", "indices")])
    work = np.full(6 + 3 * 6, 0x4545_4545, np.uint32)
    kinds = np.full(1, 0x4545_4545, np.uint32)
    oracle.coarse_count_bins(cfg, inputs, work, kinds)
    assert work[0] == sc["expected"]
    assert work[3] == 0
    assert kinds[0] == (1 if sc["expected"] == 0 else 0)
    assert np.all(work[6:] == 0x4545_4545)


# ---------------------------------------------------------------------------
# pixel math
# ---------------------------------------------------------------------------

LERP_CASES = [(6, 0, 0x3f6a_aaab, 0), (7, 1, 0x3f6a_aaab, 1), (10, 0, 0x3f59_999a, 1), (11, 0, 0x3f5d_1746, 1)]


def test_oracle_fused_lerp_half_channel_rounding_kat():
    req = np.array([[0, 0, a * 0x0101_0101, b * 0x0101_0101, 0, t, 0, 0] for a, b, t, _ in LERP_CASES], np.uint32)
    out = oracle.pixel_math_words(req)
    assert out[:, 7].tolist() == [e * 0x0101_0101 for *_, e in LERP_CASES]


def test_oracle_pixel_math_closed_forms():
    rng = np.random.default_rng(5)
    n = 4000
    a = rng.integers(0, 256, n)
    b = rng.integers(0, 256, n)
    req = np.zeros((n, 8), np.uint32)
    req[:, 0], req[:, 1] = a, b
    out = oracle.pixel_math_words(req)
    # mul_div255 is exact rounding of a*b/255
    assert np.array_equal(out[:, 0], np.floor(a * b / 255 + 0.5).astype(np.uint32))


# ---------------------------------------------------------------------------
# coverage: FP32 row sweep vs exact FP64 areas
# ---------------------------------------------------------------------------

def _poly_lines(points):
    pts = np.asarray(points, np.float32)
    n = len(pts)
    arr = np.zeros((n, 6), np.uint32)
    for i in range(n):
        arr[i, 2:4] = pts[i].view(np.uint32)
        arr[i, 4:6] = pts[(i + 1) % n].view(np.uint32)
    return arr


def test_f64_pixel_area_closed_forms():
    sq = _poly_lines([(2.25, 3.5), (12.25, 3.5), (12.25, 13.5), (2.25, 13.5)])  # y-down clockwise
    area = lambda px, py: oracle.polygon_pixel_area_f64(sq, px, py)
    assert abs(abs(area(5, 6)) - 1.0) < 1e-15
    assert abs(abs(area(2, 6)) - 0.75) < 1e-15
    assert abs(abs(area(2, 3)) - 0.375) < 1e-15
    assert abs(area(20, 6)) < 1e-15
    total = sum(area(x, y) for x in range(0, 16) for y in range(0, 16))
    assert abs(abs(total) - 100.0) < 1e-12
    tri = _poly_lines([(0.0, 0.0), (4.0, 0.0), (0.0, 4.0)])
    t = sum(oracle.polygon_pixel_area_f64(tri, x, y) for x in range(0, 5) for y in range(0, 5))
    assert abs(abs(t) - 8.0) < 1e-12
    assert abs(abs(oracle.polygon_pixel_area_f64(tri, 0, 0)) - 1.0) < 1e-15
    assert abs(abs(oracle.polygon_pixel_area_f64(tri, 3, 0)) - 0.5) < 1e-15


def _coverage_requests(a, out, limit=20000, seed=0):
    """(particle, pixel) requests for every FILL record of a frame: 8-word rows."""
    ptcl = out["ptcl"]
    fills = ptcl[(ptcl[:, 0] == 1) | (ptcl[:, 0] == 3)]
    rng = np.random.default_rng(seed)
    if len(fills) > limit // 16:
        fills = fills[rng.choice(len(fills), limit // 16, replace=False)]
    req = []
    for rec in fills:
        for _ in range(16):
            x, y = rng.integers(0, 16, 2)
            req.append([rec[3], rec[4], rec[1], x, y, rec[2], 0, 0])
    return np.array(req, np.uint32)


def test_fp32_row_sweep_within_1e5_of_fp64_area():
    a = scenes.mixed_scene_640x360().arrays()
    _, out = oracle.render_frame(a)
    req = _coverage_requests(a, out)
    segs = out["segments"]
    _, cov32 = oracle.fill_coverage_words(req, segs)
    cov64 = oracle.fill_coverage_f64(req, segs)
    err = np.abs(cov32.astype(np.float64) - cov64)
    assert err.max() <= 1e-5, err.max()


def _shuffle_within_ranges(segs, req, rng):
    """Copy of `segs` with every requested [start, end) range permuted."""
    out = segs.copy()
    for s0, s1 in {(int(r[0]), int(r[1])) for r in req}:
        out[s0:s1] = segs[s0:s1][rng.permutation(s1 - s0)]
    return out


@pytest.mark.parametrize("scene", ["mixed_scene_640x360", "clipped_paths_small"])
def test_exact_sums_are_order_free_and_within_1_lsb_of_fp32_order(scene):
    """The fine stage's coverage (three exact sums) does not depend on the slot order
    emit's atomics give a tile's segments, and differs from the reference's literal
    FP32 row sweep (shared/coverage.hlsli:34-51), in any order, by at most 1 LSB of
    alpha and a few FP32 ulps of coverage."""
    canvas = scenes.mixed_scene_640x360() if scene == "mixed_scene_640x360" else scenes.clipped_paths_small(16)
    a = canvas.arrays()
    _, out = oracle.render_frame(a)
    req = _coverage_requests(a, out, limit=40000, seed=5)
    segs = out["segments"]
    rng = np.random.default_rng(8)
    shuffled = _shuffle_within_ranges(segs, req, rng)
    alpha, cov = oracle.fill_coverage_words(req, segs)
    alpha_s, cov_s = oracle.fill_coverage_words(req, shuffled)
    assert np.array_equal(alpha, alpha_s) and np.array_equal(cov.view(np.uint32), cov_s.view(np.uint32))
    for order in (segs, shuffled):
        alpha_q, cov_q = oracle.fill_coverage_words(req, order, sequential=True)
        assert np.abs(alpha.astype(np.int64) - alpha_q).max() <= 1
        assert np.abs(cov.astype(np.float64) - cov_q).max() <= 1e-5
    # the sequential form really is order-dependent somewhere, or this test proves nothing new
    _, c0 = oracle.fill_coverage_words(req, segs, sequential=True)
    _, c1 = oracle.fill_coverage_words(req, shuffled, sequential=True)
    assert not np.array_equal(c0.view(np.uint32), c1.view(np.uint32))


def test_tile_segment_areas_reproduce_polygon_area():
    """Backdrop + clipped segments of a tile integrate to the path's exact winding area
    over that tile (up to the scan's 1e-3 px tile-edge nudges)."""
    canvas = tk.Canvas(96, 96)
    star = tk.BezPath().move_to(48, 3)
    for k in range(1, 10):
        r = 44 if k % 2 == 0 else 17
        ang = -np.pi / 2 + k * np.pi / 5
        star.line_to(48 + r * np.cos(ang), 48 + r * np.sin(ang))
    canvas.push_path(star.close_path(), tk.Color(1, 0, 0))
    a = canvas.arrays()
    _, out = oracle.render_frame(a)
    rec = a.paths[0]
    x0, y0, x1, y1 = rec[6:10]
    for ty in range(y0, y1):
        for tx in range(x0, x1):
            b = rec[4] + (ty - y0) * (x1 - x0) + (tx - x0)
            r0, r1 = out["segment_ranges"][b]
            req = np.array([[r0, r1, out["backdrops"][b], x, y, 0, 0, 0] for y in range(16) for x in range(16)],
                           np.uint32)
            tile_area = oracle.fill_coverage_f64(req, out["segments"]).sum()
            exact = sum(oracle.polygon_pixel_area_f64(a.lines, tx * 16 + x, ty * 16 + y)
                        for y in range(16) for x in range(16))
            assert abs(tile_area - exact) < 0.05, (tx, ty, tile_area, exact)


def test_single_rect_path_renders_exact_coverage():
    """An opaque axis-aligned rectangle path: interior pixels opaque, fractional edges
    quantised from their exact area, everything else clear."""
    canvas = tk.Canvas(48, 40)
    canvas.push_path(tk.BezPath().move_to(5.25, 4.5).line_to(30.75, 4.5).line_to(30.75, 21.0).line_to(5.25, 21.0)
                     .close_path(), tk.Color(1.0, 1.0, 1.0, 1.0))
    img, _ = oracle.render_frame(canvas.arrays())
    alpha = (img >> 24).astype(np.int64)
    expect = np.zeros_like(alpha)
    for y in range(40):
        for x in range(48):
            cx = max(0.0, min(x + 1, 30.75) - max(x, 5.25))
            cy = max(0.0, min(y + 1, 21.0) - max(y, 4.5))
            expect[y, x] = int(cx * cy * 255 + 0.5)
    assert np.abs(alpha - expect).max() <= 1


def test_sdf_rounded_rect_matches_f64_distance_coverage():
    canvas = tk.Canvas(64, 64)
    canvas.push_rect(8.5, 10.25, 55.0, 50.75, 9.0, tk.Color(1, 1, 1, 1))
    img, _ = oracle.render_frame(canvas.arrays())
    alpha = (img >> 24).astype(np.int64)
    x0, y0, x1, y1, r = 8.5, 10.25, 55.0, 50.75, 9.0
    cx, cy, hx, hy = (x0 + x1) / 2, (y0 + y1) / 2, (x1 - x0) / 2, (y1 - y0) / 2
    worst = 0
    for y in range(64):
        for x in range(64):
            px, py = x + 0.5 - cx, y + 0.5 - cy
            qx, qy = abs(px) - hx + r, abs(py) - hy + r
            d = min(max(qx, qy), 0.0) + np.hypot(max(qx, 0.0), max(qy, 0.0)) - r
            cov = min(max(0.5 - d, 0.0), 1.0)
            worst = max(worst, abs(alpha[y, x] - int(cov * 255 + 0.5)))
    assert worst <= 1


def test_oracle_results_do_not_depend_on_threads():
    # Threads split lines by path, chunks and tiles; every output is identical to one
    # thread's, the emitted segment order included.
    arrays = scenes.clipped_paths_small(8).arrays()
    runs = []
    for n in (1, 3, 8):
        oracle.set_threads(n)
        runs.append(oracle.render_frame(arrays))
    oracle.set_threads(min(8, len(os.sched_getaffinity(0))))
    (image, out) = runs[0]
    for other_image, other in runs[1:]:
        assert np.array_equal(image, other_image)
        for key in out:
            if key != "stage_ms":
                assert np.array_equal(out[key], other[key]), key
