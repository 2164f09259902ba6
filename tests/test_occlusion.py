"""Early occlusion of tile bins: an opaque solid SDF rectangle hides the earlier
draws of its DrawBatch in the whole tiles it covers, before coarse.

The rule is TileDrawBins::reset_transient with Occlusion::update
(src/shared/gpu_plan/tile_bins/transient.rs:7-146, occlusion.rs:48-160,
occlusion/geometry.rs:17-135, occlusion/transient.rs:9-66), the path an
immediate canvas takes.  The recorder (canvas.cpp) builds the bins the device
reads; the oracle restates the rule on its own.  The tests here hold both to:

- the reference's own expectations for immediate canvases
  (src/shared/gpu_plan/tile_bins/occlusion/tests.rs: the tests that build
  transient bins);
- the closed form of the proof region, in exact rationals: a tile is culled only
  if its square lies inside the rectangle's central part (corners and the
  fractional guard excluded) and inside every clip of its batch;
- invariance: culling never changes the oracle's image.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import scenes

BLACK = tk.Color(0, 0, 0, 1)
WHITE = tk.Color(1, 1, 1, 1)


def rgba8(r, g, b, a):
    return tk.Color(r / 255, g / 255, b / 255, a / 255)


def full(canvas, brush, radius=0.0):
    return canvas.push_rect(0, 0, canvas.width, canvas.height, radius, brush)


def tile_lists(tile_draws, indices):
    out = []
    for start, n in np.asarray(tile_draws).reshape(-1, 2):
        out.append([] if n == 0 else [int(v) for v in indices[start:start + n]])
    return out


def bins(canvas):
    """Recorder bins, checked equal to the oracle's; returns per-tile lists."""
    arrays = canvas.arrays(0)
    want_td, want_idx = oracle.tile_draw_bins(arrays)
    assert np.array_equal(arrays.tile_draws, want_td)
    assert np.array_equal(arrays.tile_draw_indices, want_idx)
    return tile_lists(arrays.tile_draws, arrays.tile_draw_indices)


def unculled(arrays):
    return tile_lists(*oracle.tile_draw_bins(arrays, cull=False))


def scene(alpha, radius=0.0):
    # occlusion/tests.rs:11-19: an opaque background and a full-canvas rect of `alpha`.
    canvas = tk.Canvas(64, 64)
    full(canvas, BLACK)
    full(canvas, rgba8(23, 47, 93, alpha), radius)
    return canvas


@pytest.mark.parametrize("count", [1, 255, 256, 257, 513])
def test_hidden_references_are_removed(count):
    # early_occlusion_removes_hidden_gpu_refs_but_keeps_filter_source_queries (tests.rs:37-60)
    canvas = scene(120)
    for _ in range(count):
        full(canvas, WHITE)
    lists = bins(canvas)
    assert sum(len(v) for v in lists) == 16
    assert all(v == [count + 1] for v in lists)


@pytest.mark.parametrize("alpha", [0, 120, 254])
def test_translucent_rect_proves_nothing(alpha):
    # early_occlusion_preserves_alpha_aa_corners_and_uncertain_transforms (tests.rs:62-84)
    lists = bins(scene(alpha))
    assert lists[0] == [0, 1]
    assert lists == unculled(scene(alpha).arrays(0))


def test_rounded_corners_keep_their_tiles():
    lists = bins(scene(255, 16.0))
    assert lists[0] == [0, 1]
    assert lists[5] == [1]


@pytest.mark.parametrize("transform", [(1, 0, 0, 1, 0.375, 0.375), (0.9, 0, 0, 0.9, 0, 0),
                                       (np.cos(0.1), np.sin(0.1), -np.sin(0.1), np.cos(0.1), 0, 0)])
def test_uncertain_transforms_prove_nothing(transform):
    # The same test's translate(0.375) / scale(0.9) / rotate(0.1) of the occluder.  The
    # recorder bakes transforms into SDF coordinates, so the draw record's transform
    # is set on the arrays and only the oracle's rule sees it.
    arrays = scene(255).arrays(0)
    draws = np.array(arrays.draws, copy=True)
    draws[1, 19:25] = np.asarray(transform, np.float32).view(np.uint32)
    arrays.draws = draws
    assert tile_lists(*oracle.tile_draw_bins(arrays))[0] == [0, 1]


def test_every_draw_survives_at_partial_clip_edges():
    # early_occlusion_preserves_every_draw_at_partial_clip_edges (tests.rs:86-98)
    canvas = tk.Canvas(64, 64)
    canvas.push_clip_sdf_rect_layer(4.375, 4.375, 59.625, 59.625, 8.0)
    full(canvas, rgba8(200, 40, 90, 119))
    full(canvas, WHITE)
    canvas.pop_layer()
    lists = bins(canvas)
    assert lists[0] == [1, 2]
    assert lists[5] == [2]


def test_occlusion_stays_inside_each_batch():
    # early_occlusion_stays_inside_each_batch_and_keeps_fused_wrappers (tests.rs:100-124),
    # the clip-layer case (opacity and blend layers are not recorded on this path).
    canvas = scene(120)
    canvas.push_clip_sdf_rect_layer(4, 4, 60, 60, 8.0)
    full(canvas, BLACK)
    full(canvas, WHITE)
    canvas.pop_layer()
    full(canvas, rgba8(90, 80, 70, 120))
    assert bins(canvas)[5] == [0, 1, 4, 5]


def test_unknown_brushes_and_path_clips_prove_nothing():
    # early_occlusion_rejects_unknown_brush_and_clip_coverage (tests.rs:245-266)
    canvas = tk.Canvas(64, 64)
    full(canvas, BLACK)
    full(canvas, tk.FourCornerGradient((0, 0, 64, 64), [WHITE] * 4))
    assert bins(canvas) == unculled(canvas.arrays(0))
    canvas = tk.Canvas(64, 64)
    square = tk.BezPath().move_to(0, 0).line_to(64, 0).line_to(64, 64).line_to(0, 64).close_path()
    canvas.push_clip_layer(square, transform=(1, 0, 0, 1, 0.25, 0.25))
    full(canvas, BLACK)
    full(canvas, WHITE)
    canvas.pop_layer()
    assert bins(canvas) == unculled(canvas.arrays(0))


def test_largest_proof_of_a_batch_is_kept():
    # occlusion/transient.rs:9-66: walking back, a proof replaces the current one only
    # when it covers strictly more tiles (never a union), so a later, larger rectangle
    # keeps a smaller earlier one from hiding anything outside the larger one.
    canvas = tk.Canvas(64, 32)
    full(canvas, BLACK)                         # 0
    full(canvas, rgba8(90, 80, 70, 120))        # 1
    canvas.push_rect(0, 0, 32, 32, 0.0, WHITE)  # 2: proof x 0..2
    full(canvas, rgba8(90, 80, 70, 100))        # 3
    canvas.push_rect(32, 0, 64, 32, 0.0, WHITE)  # 4: proof x 2..4, same size
    full(canvas, rgba8(90, 80, 70, 50))         # 5
    lists = bins(canvas)
    assert lists[2] == [4, 5]
    assert lists[0] == [0, 1, 2, 3, 5]  # draw 2's proof is not larger: nothing hidden there


def central_tiles(x0, y0, x1, y1, radius, tiles_w, tiles_h):
    """Tiles whose square lies inside the rect's central part, in exact rationals."""
    v = [np.float32(x0), np.float32(y0), np.float32(x1), np.float32(y1)]
    r = np.float32(radius)
    whole = all(float(t) == int(t) for t in v + [r])
    guard = Fraction(0) if whole else Fraction(1, 32)
    fx0, fy0, fx1, fy1 = (Fraction(float(t)) for t in v)
    fr = min(max(Fraction(float(r)), Fraction(0)), max(min(fx1 - fx0, fy1 - fy0) / 2, Fraction(0)))
    bx0, by0 = max(int(np.floor(float(v[0]))), 0) // 16, max(int(np.floor(float(v[1]))), 0) // 16
    bx1 = min((max(int(np.ceil(float(v[2]))), 0) + 15) // 16, tiles_w)
    by1 = min((max(int(np.ceil(float(v[3]))), 0) + 15) // 16, tiles_h)
    return {(tx, ty) for ty in range(by0, by1) for tx in range(bx0, bx1)
            if 16 * tx >= fx0 + fr + guard and 16 * (tx + 1) <= fx1 - fr - guard
            and 16 * ty >= fy0 + fr + guard and 16 * (ty + 1) <= fy1 - fr - guard}


@pytest.mark.parametrize("seed", range(40))
def test_culled_tiles_match_the_closed_form(seed):
    rng = np.random.default_rng(seed)
    w, h = 160, 128
    canvas = tk.Canvas(w, h)
    full(canvas, BLACK)
    full(canvas, rgba8(10, 200, 30, 100))
    if seed % 2:
        x0, y0 = rng.integers(-8, 60, 2)
        x1, y1 = x0 + rng.integers(20, 140), y0 + rng.integers(20, 120)
        radius = float(rng.integers(0, 12))
    else:
        x0, y0 = rng.uniform(-8, 60, 2)
        x1, y1 = x0 + rng.uniform(20, 140), y0 + rng.uniform(20, 120)
        radius = float(rng.uniform(0, 12))
    canvas.push_rect(x0, y0, x1, y1, radius, WHITE)
    full(canvas, rgba8(90, 80, 70, 50))
    lists = bins(canvas)
    tw, th = (w + 15) // 16, (h + 15) // 16
    culled = {(t % tw, t // tw) for t, v in enumerate(lists) if v[:1] == [2]}
    assert culled == central_tiles(x0, y0, x1, y1, radius, tw, th)
    plain = unculled(canvas.arrays(0))
    for t, v in enumerate(lists):
        assert v == ([2, 3] if (t % tw, t // tw) in culled else plain[t])


def test_c1_counts_and_invariant_image():
    canvas = scenes.mixed_scene_640x360()
    arrays = canvas.arrays(0)
    plain_td, plain_idx = oracle.tile_draw_bins(arrays, cull=False)
    # Two opaque solid rects (draws 4 and 207) carry proofs of 6 and 14 tiles; the
    # larger, later one is kept over the whole single batch, so draw 4's is unused.
    lists = bins(canvas)
    plain = tile_lists(plain_td, plain_idx)
    changed = [t for t in range(len(lists)) if lists[t] != plain[t]]
    assert len(changed) == 14
    for t in changed:
        k = plain[t].index(207)
        assert lists[t] == plain[t][k:]
    assert len(plain_idx) - arrays.tile_draw_index_count == 1303
    assert len(plain_idx) == 197209
    # Culled references are hidden anyway: the image is unchanged.
    image, _ = oracle.render_frame(arrays)
    arrays.tile_draws, arrays.tile_draw_indices = plain_td, plain_idx
    arrays.scalars["tile_draw_index_count"] = len(plain_idx)
    image_plain, _ = oracle.render_frame(arrays)
    assert np.array_equal(image, image_plain)


@pytest.mark.parametrize("seed", range(6))
def test_random_occluded_scenes_render_unchanged(seed):
    rng = np.random.default_rng(100 + seed)
    canvas = tk.Canvas(200, 150)
    depth = 0
    for k in range(60):
        u = rng.random()
        if u < 0.08 and depth < 3:
            x0, y0 = rng.uniform(0, 80, 2)
            canvas.push_clip_sdf_rect_layer(x0, y0, x0 + rng.uniform(40, 140), y0 + rng.uniform(40, 120),
                                            float(rng.uniform(0, 10)))
            depth += 1
        elif u < 0.14 and depth:
            canvas.pop_layer()
            depth -= 1
        else:
            x0, y0 = rng.uniform(-10, 180, 2)
            alpha = 1.0 if rng.random() < 0.4 else rng.uniform(0.2, 1)
            color = tk.Color(*rng.uniform(0, 1, 3), alpha)
            if rng.random() < 0.6:
                canvas.push_rect(x0, y0, x0 + rng.uniform(8, 120), y0 + rng.uniform(8, 100),
                                 float(rng.integers(0, 3) * rng.uniform(0, 12)), color)
            else:
                p = tk.BezPath().move_to(x0, y0)
                for _ in range(3):
                    p.curve_to(*rng.uniform(0, 200, 6))
                canvas.push_path(p.close_path(), color)
    while depth:
        canvas.pop_layer()
        depth -= 1
    arrays = canvas.arrays(0)
    bins(canvas)
    image, _ = oracle.render_frame(arrays)
    plain_td, plain_idx = oracle.tile_draw_bins(arrays, cull=False)
    arrays.tile_draws, arrays.tile_draw_indices = plain_td, plain_idx
    arrays.scalars["tile_draw_index_count"] = len(plain_idx)
    image_plain, _ = oracle.render_frame(arrays)
    assert np.array_equal(image, image_plain)
