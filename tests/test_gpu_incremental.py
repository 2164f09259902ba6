"""Incremental rendering over a damaged-tile list (tk_frame_render_tiles): the
active_tile_list dispatch of scan, coarse and fine (scan/index.hlsli:5-8,
coarse/active_tiles.hlsli:6-9, fine/main.hlsl:26-27).  After a scene change
confined to some tiles, re-rendering just those tiles must give the image a full
render of the changed scene gives, bit for bit (the oracle renders the changed
scene in full)."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import _abi, scenes

pytestmark = pytest.mark.gpu


def copy_arrays(arrays, **changes):
    out = tk.SceneArrays.__new__(tk.SceneArrays)
    out.__dict__.update(arrays.__dict__)
    for name, value in changes.items():
        setattr(out, name, value)
    return out


def draw_tiles(arrays, draw):
    """The tiles of a draw's pixel bounds (every tile a draw can change)."""
    x0, y0, x1, y1 = (int(v) for v in arrays.draws[draw, 10:14].view(np.int32))
    tw, th = arrays.tiles_width, arrays.tiles_height
    tx0, ty0 = max(x0, 0) // 16, max(y0, 0) // 16
    tx1, ty1 = min((max(x1, 0) + 15) // 16, tw), min((max(y1, 0) + 15) // 16, th)
    return [ty * tw + tx for ty in range(ty0, ty1) for tx in range(tx0, tx1)]


def solid_draws(arrays):
    base = arrays.paint_brush_base
    return [d for d in range(len(arrays.draws))
            if arrays.draws[d, 8] == 0 and arrays.paint[base + arrays.draws[d, 6]] == _abi.BRUSH_SOLID]


def test_all_tiles_in_any_order_equal_a_full_render(ctx):
    arrays = scenes.clipped_paths_small(16).arrays()
    scene = arrays.to_scene()
    full = tk.Frame(ctx, scene)
    full.render()
    expected = full.image()
    frame = tk.Frame(ctx, scene)
    T = arrays.tiles_width * arrays.tiles_height
    frame.render_tiles(np.random.default_rng(3).permutation(T))
    assert np.array_equal(frame.image(), expected)
    assert frame.stats()["overflow"] == 0


def test_repainting_any_tiles_of_an_unchanged_scene_keeps_the_image(ctx):
    arrays = scenes.clipped_paths_small(16).arrays()
    frame = tk.Frame(ctx, arrays.to_scene())
    frame.render()
    before = frame.image()
    T = arrays.tiles_width * arrays.tiles_height
    rng = np.random.default_rng(4)
    for count in (1, 7, T // 3):
        frame.render_tiles(rng.choice(T, size=count, replace=False))
        assert np.array_equal(frame.image(), before), count
    frame.render_tiles(np.array([5, 5, 5, 2, 5], np.uint32))  # duplicates are ignored
    assert np.array_equal(frame.image(), before)


def test_brush_change_rerendered_on_its_tiles_matches_full_render(ctx):
    arrays = scenes.clipped_paths_small(16).arrays()
    frame = tk.Frame(ctx, arrays.to_scene())
    frame.render()
    changed = [d for d in solid_draws(arrays) if len(draw_tiles(arrays, d)) <= 12][::7][:6]
    assert len(changed) >= 3
    paint = arrays.paint.copy()
    for k, d in enumerate(changed):
        # opaque and translucent colours both, so occlusion skips change too
        paint[arrays.paint_brush_base + arrays.draws[d, 6] + 4] = [0xFF20C040, 0x80104020][k % 2]
    new = copy_arrays(arrays, paint=paint)
    damage = sorted({t for d in changed for t in draw_tiles(new, d)})
    frame.update(new.to_scene())
    frame.render_tiles(damage)
    image, _ = oracle.render_frame(new)
    got = frame.image()
    assert not np.array_equal(got, oracle.render_frame(arrays)[0])  # the change is visible
    assert np.array_equal(got, image)


def test_clip_change_rerendered_on_its_tiles_matches_full_render(ctx):
    arrays = scenes.clipped_paths_small(16).arrays()
    frame = tk.Frame(ctx, arrays.to_scene())
    frame.render()
    # The smallest clip rectangle wider and taller than 40 px shrinks by 9 px on every side.
    def rect_of(draw):
        sdf = int(arrays.draws[draw, 2])
        return arrays.paint[sdf + 1:sdf + 5].view(np.float32)
    clips = [d for d in sorted({int(v) for v in arrays.layers[:, 1]})
             if arrays.paint[int(arrays.draws[d, 2])] == 1 and min(rect_of(d)[2:] - rect_of(d)[:2]) > 40]
    clip = min(clips, key=lambda d: len(draw_tiles(arrays, d)))
    sdf = int(arrays.draws[clip, 2])
    paint = arrays.paint.copy()
    rect = paint[sdf + 1:sdf + 5].view(np.float32)
    assert paint[sdf] == 1 and rect[2] - rect[0] > 40 and rect[3] - rect[1] > 40
    rect[0:2] += 9.0
    rect[2:4] -= 9.0
    new = copy_arrays(arrays, paint=paint)
    frame.update(new.to_scene())
    frame.render_tiles(draw_tiles(new, clip))
    image, _ = oracle.render_frame(new)
    assert np.array_equal(frame.image(), image)


def test_render_tiles_rejects_out_of_range_and_ignores_an_empty_list(ctx):
    arrays = scenes.mixed_scene_640x360().arrays()
    frame = tk.Frame(ctx, arrays.to_scene())
    frame.render()
    before = frame.image()
    T = arrays.tiles_width * arrays.tiles_height
    with pytest.raises(tk.TileinkError) as err:
        frame.render_tiles([0, T])
    assert err.value.status == _abi.TK_ERR_INVALID
    frame.render_tiles([])
    assert np.array_equal(frame.image(), before)
    frame.render()
    assert np.array_equal(frame.image(), before)
