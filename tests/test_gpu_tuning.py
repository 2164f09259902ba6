"""Launch parameters (tk_frame_set_tuning, tk_frame_autotune): the XCD run
lengths only renumber workgroups, so every value must give the same image and
tables, for a single canvas, an incremental render and a batch of scenes."""
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


def test_autotune_picks_valid_values_and_keeps_the_image(ctx):
    frame = tk.Frame(ctx, scenes.clipped_paths_small(16).encode(0))
    frame.render()
    ref = frame.image()
    chosen = frame.autotune(repeats=1)
    assert chosen == frame.tuning()
    for v in chosen.values():
        assert v == 0 or (v & (v - 1)) == 0
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
    assert frame.tuning() == before
