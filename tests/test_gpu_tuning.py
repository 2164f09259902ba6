"""Launch parameters (tk_frame_set_tuning, tk_frame_autotune): XCD run lengths,
persistent launches over the per-XCD work queues, scan and coarse workgroup
sizes and the coarse scan's block size only change how work is dealt, so every
value must give the same image and tables, for a single canvas, an incremental
render and a batch of scenes."""
import numpy as np
import pytest

import tileink_amd as tk
from tileink_amd import _abi, scenes

pytestmark = pytest.mark.gpu

RUNS = [(0, 0, 0), (16, 16, 32), (64, 128, 512), (1, 1, 1), (1024, 1024, 1024)]


def render_with(frame, tuning):
    frame.set_tuning(*tuning)
    frame.render()
    return frame.image()


def test_every_run_length_renders_the_same_frame(ctx):
    canvas = scenes.clipped_paths_small(16)
    frame = tk.Frame(ctx, canvas.encode(0))
    ref = render_with(frame, (0, 64, 256))
    kinds = frame.buffer(_abi.BUF_TILE_KINDS, np.uint32, frame.batch_count * frame.tiles)
    for t in RUNS:
        assert np.array_equal(render_with(frame, t), ref), t
        assert np.array_equal(frame.buffer(_abi.BUF_TILE_KINDS, np.uint32, frame.batch_count * frame.tiles), kinds), t
    assert frame.stats()["overflow"] == 0


def test_run_lengths_apply_to_incremental_renders_and_batches(ctx):
    arrays = scenes.clipped_paths_small(16).arrays()
    frame = tk.Frame(ctx, arrays.to_scene())
    frame.render()
    full = frame.image()
    T = arrays.tiles_width * arrays.tiles_height
    frame.set_tuning(16, 16, 32)
    frame.render_tiles(np.random.default_rng(1).permutation(T)[: T // 3])
    assert np.array_equal(frame.image(), full)

    canvases = [scenes.scene_set_member(w, h, n, i) for i, (w, h, n) in enumerate([(96, 64, 40), (300, 140, 120)])]
    batch = tk.Batch(ctx, [c.encode(0) for c in canvases])
    batch.render()
    ref = [batch.image(i) for i in range(len(canvases))]
    for t in RUNS:
        batch.frame.set_tuning(*t)
        batch.render()
        assert all(np.array_equal(batch.image(i), ref[i]) for i in range(len(canvases))), t


# (persistent, scan_block, coarse_block, coarse_scan_items) beside the default (0, 256, 4, 4096)
SHAPES = [(1, 64, 1, 1024), (2, 1024, 2, 8192), (3, 128, 4, 2048), (3, 512, 1, 4096), (2, 256, 4, 1024)]


def shaped(frame, shape, runs=(0, 64, 256)):
    p, sb, cb, items = shape
    frame.set_tuning(*runs, persistent=p, scan_block=sb, coarse_block=cb, coarse_scan_items=items)


def frame_tables(frame):
    QT = frame.batch_count * frame.tiles
    K = int(frame.stats()["ptcl_records"])
    return (frame.image(), frame.buffer(_abi.BUF_TILE_KINDS, np.uint32, QT),
            frame.buffer(_abi.BUF_TILE_PTCL_RANGES, np.uint32, 2 * QT), frame.buffer(_abi.BUF_PTCL, np.uint32, 6 * K))


@pytest.mark.parametrize("scene", ["clipped", "mixed"])
def test_persistent_launches_and_block_sizes_render_the_same_frame(ctx, scene):
    canvas = scenes.clipped_paths_small(8) if scene == "clipped" else scenes.mixed_scene_640x360()
    frame = tk.Frame(ctx, canvas.encode(0))
    frame.render()
    ref = frame_tables(frame)
    for shape in SHAPES:
        for runs in ((0, 64, 256), (16, 0, 0), (64, 1, 32)):
            shaped(frame, shape, runs)
            frame.render()
            got = frame_tables(frame)
            for name, a, b in zip(("image", "kinds", "ranges", "ptcl"), got, ref):
                assert np.array_equal(a, b), (shape, runs, name)
            assert frame.stats()["overflow"] == 0
    # the persistent launches replay from a captured graph (their counters are
    # cleared by a fill kernel inside it)
    shaped(frame, (3, 256, 4, 4096))
    frame.set_graph(True)
    for _ in range(3):
        frame.render()
        assert np.array_equal(frame.image(), ref[0])


def test_persistent_launches_in_incremental_renders_and_batches(ctx):
    arrays = scenes.clipped_paths_small(16).arrays()
    frame = tk.Frame(ctx, arrays.to_scene())
    frame.render()
    full = frame.image()
    T = arrays.tiles_width * arrays.tiles_height
    shaped(frame, (3, 128, 2, 2048))
    frame.render_tiles(np.random.default_rng(2).permutation(T)[: T // 2])
    assert np.array_equal(frame.image(), full)
    canvases = [scenes.scene_set_member(w, h, n, i) for i, (w, h, n) in
                enumerate([(96, 64, 40), (300, 140, 120), (700, 333, 900), (40, 1000, 60)])]
    batch = tk.Batch(ctx, [c.encode(0) for c in canvases])
    batch.render()
    ref = [batch.image(i) for i in range(len(canvases))]
    for shape in SHAPES:
        shaped(batch.frame, shape)
        batch.render()
        assert all(np.array_equal(batch.image(i), ref[i]) for i in range(len(canvases))), shape


def test_autotune_picks_valid_values_and_keeps_the_image(ctx):
    frame = tk.Frame(ctx, scenes.clipped_paths_small(16).encode(0))
    frame.render()
    ref = frame.image()
    chosen = frame.autotune(repeats=1)
    assert chosen == frame.tuning()
    for k in ("scan_xcd_run", "coarse_xcd_run", "fine_xcd_run"):
        assert chosen[k] == 0 or (chosen[k] & (chosen[k] - 1)) == 0
    assert chosen["persistent"] in (0, 1, 2, 3)
    assert chosen["scan_block"] in (64, 128, 256, 512, 1024)
    assert chosen["coarse_block"] in (1, 2, 4)
    assert chosen["coarse_scan_items"] in (1024, 2048, 4096, 8192)
    assert np.array_equal(frame.image(), ref)
    frame.render()
    assert np.array_equal(frame.image(), ref)


def test_invalid_run_lengths_are_rejected(ctx):
    frame = tk.Frame(ctx, scenes.mixed_scene_640x360().encode(0))
    before = frame.tuning()
    for bad in ((3, 0, 0), (0, 2048, 0), (0, 0, 100)):
        with pytest.raises(tk.TileinkError) as err:
            frame.set_tuning(*bad)
        assert err.value.status == _abi.TK_ERR_INVALID
    for bad in (dict(persistent=4), dict(scan_block=32), dict(scan_block=384), dict(coarse_block=3),
                dict(coarse_block=8), dict(coarse_scan_items=512), dict(coarse_scan_items=3000),
                dict(diff_keep_coverage=2)):
        with pytest.raises(tk.TileinkError) as err:
            frame.set_tuning(**bad)
        assert err.value.status == _abi.TK_ERR_INVALID, bad
    assert frame.tuning() == before
