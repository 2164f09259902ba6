"""Batches of independent scenes rendered as one frame (tk_batch_*): every scene's
image equals the image of its own frame (itself checked against the oracle by the
frame tests), for scenes of different sizes, clear colours, clip nesting and
brushes; repeated and graph-replayed batch renders give the same images."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import scenes

from test_gpu_brushes import gradient_canvas

pytestmark = pytest.mark.gpu


def batch_cases():
    canvases = [scenes.mixed_scene_640x360(), scenes.scene_set_member(100, 37, 40, 0), scenes.scene_set_member(333, 250, 300, 1),
                tk.Canvas(17, 9), scenes.clipped_paths_small(16), gradient_canvas(["linear", "radial", "sweep", "four_corner"])]
    clears = [0, 0x80402010, 0, 0xFF0000FF, 0, 0x20202020]
    return canvases, [c.encode(clear) for c, clear in zip(canvases, clears)]


def test_batch_images_equal_single_scene_renders(ctx):
    canvases, encoded = batch_cases()
    batch = tk.Batch(ctx, encoded)
    batch.render()
    assert batch.frame.stats()["overflow"] == 0
    for i, scene in enumerate(encoded):
        single = tk.Frame(ctx, scene)
        single.render()
        expected = single.image()
        got = batch.image(i)
        assert got.shape == (scene.height, scene.width)
        assert np.array_equal(got, expected), (i, int((got != expected).sum()))
    assert np.all(batch.image(3) == 0xFF0000FF)  # the empty canvas keeps its clear colour
    first = [batch.image(i) for i in range(len(batch))]
    batch.render()
    assert all(np.array_equal(batch.image(i), first[i]) for i in range(len(batch)))
    batch.frame.set_graph(True)
    for _ in range(3):
        batch.render()
        assert all(np.array_equal(batch.image(i), first[i]) for i in range(len(batch)))
    assert np.array_equal(first[0], oracle.render_frame(canvases[0].arrays(0))[0])


def test_batch_rejects_empty_batches_and_incremental_renders(ctx):
    with pytest.raises(tk.TileinkError):
        tk.Batch(ctx, [])
    _, encoded = batch_cases()
    batch = tk.Batch(ctx, encoded[:2])
    with pytest.raises(tk.TileinkError):
        batch.frame.render_tiles([0])
    with pytest.raises(tk.TileinkError):
        batch.image(2)
