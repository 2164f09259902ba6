"""Full renders replayed from a captured hipGraph (tk_frame_set_graph) against
eager renders: every replay, with eager work, incremental renders and
synchronisations between the replays, and after a scene update (which drops the
graph so that it is captured again)."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import scenes

from test_gpu_incremental import copy_arrays, solid_draws

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager_renders(ctx):
    arrays = scenes.clipped_paths_small(16).arrays()
    scene = arrays.to_scene()
    eager = tk.Frame(ctx, scene)
    eager.render()
    expected = eager.image()
    assert np.array_equal(expected, oracle.render_frame(arrays)[0])
    frame = tk.Frame(ctx, scene)
    frame.set_graph(True)
    for i in range(5):
        frame.render()
        assert np.array_equal(frame.image(), expected), i
        eager.render()  # eager work on the same context between replays
        assert np.array_equal(eager.image(), expected), i
    frame.render_tiles([0, 1, 17])  # an incremental render between replays
    assert np.array_equal(frame.image(), expected)
    frame.render()
    frame.render()
    assert np.array_equal(frame.image(), expected)
    assert frame.stats()["overflow"] == 0

    paint = arrays.paint.copy()
    for d in solid_draws(arrays)[::5]:
        paint[arrays.paint_brush_base + arrays.draws[d, 6] + 4] = 0xFF3060A0
    new = copy_arrays(arrays, paint=paint)
    frame.update(new.to_scene())
    changed = oracle.render_frame(new)[0]
    assert not np.array_equal(changed, expected)
    for i in range(3):
        frame.render()
        assert np.array_equal(frame.image(), changed), i
    frame.set_graph(False)
    frame.render()
    assert np.array_equal(frame.image(), changed)
