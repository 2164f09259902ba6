"""Whole-frame parity of the gfx950 pipeline against the CPU oracle (bit-exact
integer stages, canonical segment order, identical PTCL streams, image)."""
import numpy as np
import pytest

import tileink_amd as tk
from tileink_amd import scenes

import oracle

from helpers import channel_diff, compare_frame, read_frame

pytestmark = pytest.mark.gpu


def _report(gpu, ref, bad):
    return f"mismatched stages {bad}; image max channel diff {channel_diff(gpu['image'], ref['image'])}"


def test_mixed_640x360_frame_matches_oracle(ctx):
    gpu, ref, bad = compare_frame(ctx, scenes.mixed_scene_640x360())
    assert not bad, _report(gpu, ref, bad)


def test_full_size_4096px_200k_path_frame_matches_oracle(ctx):
    # The benchmark frame itself (4096x4096, 200,000 paths, 32 nested clips), every
    # stage against the oracle run threaded on the host cores.
    import os
    oracle.set_threads(min(16, len(os.sched_getaffinity(0))))
    canvas = scenes.clipped_paths_scene(4096, 4096, 200_000, clips=32, workload=2)
    print("full-size frame: scene generated", flush=True)
    gpu, ref, bad = compare_frame(ctx, canvas, hashed_segments=True)
    oracle.set_threads(min(8, len(os.sched_getaffinity(0))))
    assert not bad, _report(gpu, ref, bad)
    assert gpu["stats"]["ptcl_records"] == len(ref["ptcl"]) > 3_000_000


@pytest.mark.parametrize("scale", [16, 8, 4])
def test_small_clipped_path_frames_match_oracle(ctx, scale):
    gpu, ref, bad = compare_frame(ctx, scenes.clipped_paths_small(scale))
    assert not bad, _report(gpu, ref, bad)


def test_clear_color_and_empty_canvas(ctx):
    canvas = tk.Canvas(37, 21)
    gpu, ref, bad = compare_frame(ctx, canvas, clear=0x80402010)
    assert not bad, _report(gpu, ref, bad)
    assert np.all(gpu["image"] == 0x80402010)


def test_deep_clip_stack_spills_to_lds(ctx):
    canvas = tk.Canvas(300, 200)
    for k in range(12):
        canvas.push_clip_sdf_rect_layer(4 * k + 3.3, 3 * k + 1.7, 297 - 5 * k, 198 - 4 * k, 6.0 + k)
        path = tk.BezPath().move_to(10 + 9 * k, 5).curve_to(290, 40 + k, 10, 160, 280 - 7 * k, 195).close_path()
        canvas.push_path(path, tk.Color(0.1 * (k % 10), 0.5, 1.0 - 0.07 * k, 0.55), rule=tk.FillRule(k % 2))
    for k in range(12):
        canvas.push_rect(20 + 10 * k, 20, 60 + 10 * k, 180, 3.0, tk.Color(1, 0.2, 0.1 * k % 1, 0.3))
        canvas.pop_layer()
    gpu, ref, bad = compare_frame(ctx, canvas)
    assert not bad, _report(gpu, ref, bad)


def _frame_matches_oracle(frame, scene, arrays):
    frame.set_capture(True)
    frame.render()
    gpu = read_frame(frame, scene)
    image, ref = oracle.render_frame(arrays)
    for key in ("tile_ptcl_counts", "tile_kinds", "ptcl"):
        assert np.array_equal(gpu[key], ref[key]), key
    assert np.array_equal(gpu["image"], image), channel_diff(gpu["image"], image)


def test_frame_update_with_more_layer_pairs_than_its_class_table(ctx):
    """A frame sizes its per-tile layer class table for the distinct (draw, tag)
    pairs of the scene it was created with; an update that brings more pairs
    falls back to evaluating each layer, with the same records and image."""
    arrays = scenes.clipped_paths_small(16).arrays()
    assert len({tuple(r) for r in arrays.layers[:, :2].tolist()}) > 1
    collapsed = tk.SceneArrays.__new__(tk.SceneArrays)
    collapsed.__dict__.update(arrays.__dict__)
    collapsed.layers = arrays.layers.copy()
    collapsed.layers[:, 1] = arrays.layers[0, 1]  # every stack level clips by the first clip draw
    first = collapsed.to_scene()
    frame = tk.Frame(ctx, first)
    _frame_matches_oracle(frame, first, collapsed)
    second = arrays.to_scene()
    frame.update(second)
    _frame_matches_oracle(frame, second, arrays)
    frame.update(first)
    _frame_matches_oracle(frame, first, collapsed)


def test_full_size_frame_is_deterministic_and_matches_its_incremental_form(ctx):
    """At the benchmark's full size (4096x4096, 200 000 paths, 32 nested clips), where
    the oracle would take minutes: two renders give the same image although emit's
    segment order comes from atomics (the coverage sums are exact and order-free),
    and re-rendering every tile through the incremental path reproduces it."""
    canvas = scenes.clipped_paths_scene(4096, 4096, 200_000, clips=32, workload=2)
    scene = canvas.encode(0)  # points into the canvas's arrays: the canvas must outlive it
    frame = tk.Frame(ctx, scene)
    frame.render()
    first = frame.image()
    stats = frame.stats()
    assert stats["overflow"] == 0 and stats["ptcl_records"] > 3_000_000
    frame.render()
    assert np.array_equal(frame.image(), first)
    T = scene.tiles_width * scene.tiles_height
    frame.render_tiles(np.arange(T, dtype=np.uint32)[::-1].copy())
    assert np.array_equal(frame.image(), first)
    assert len(np.unique(first)) > 10_000


def test_update_follows_a_new_paint_split(ctx):
    # Two encodings of one scene with equal word counts but different paint
    # sections: A = sdf | brush | 3 unused words, B = sdf | 3 shadow words | brush.
    # A frame created from A and updated to B must read B's brush section.
    canvas = scenes.mixed_scene_640x360()
    base = canvas.arrays(0)
    s0, b0 = base.paint_sdf_shadow_base, base.paint_brush_base
    assert s0 == b0  # the recorder writes no shadow words
    pad = np.array([0xDEADBEEF, 0x01020304, 0xFFFFFFFF], np.uint32)
    a = tk.SceneArrays.__new__(tk.SceneArrays)
    a.__dict__.update(base.__dict__)
    a.scalars = dict(base.scalars)
    a.paint = np.concatenate([base.paint, pad])
    a.scalars["paint_words"] = len(a.paint)
    b = tk.SceneArrays.__new__(tk.SceneArrays)
    b.__dict__.update(a.__dict__)
    b.scalars = dict(a.scalars)
    b.paint = np.concatenate([base.paint[:s0], pad, base.paint[s0:]])
    b.scalars["paint_brush_base"] = s0 + len(pad)
    expected, _ = oracle.render_frame(b)
    assert np.array_equal(expected, oracle.render_frame(a)[0])
    frame = tk.Frame(ctx, a.to_scene())
    frame.render()
    assert np.array_equal(frame.image(), expected)
    frame.update(b.to_scene())
    frame.render()
    assert np.array_equal(frame.image(), expected)
    assert frame.stats()["overflow"] == 0


def test_update_rejects_paint_bases_out_of_order(ctx):
    base = scenes.mixed_scene_640x360().arrays(0)
    frame = tk.Frame(ctx, base.to_scene())
    bad = tk.SceneArrays.__new__(tk.SceneArrays)
    bad.__dict__.update(base.__dict__)
    bad.scalars = dict(base.scalars)
    bad.scalars["paint_sdf_shadow_base"] = base.paint_brush_base + 1
    with pytest.raises(tk.TileinkError):
        frame.update(bad.to_scene())
