"""tk_coarse_no_alloc (coarse from the fixed per-device scratch array) against
tk_coarse on the same device inputs: every output word equal, on two calls in a
row (the second on a smaller scene, so it reads scratch the first one left
behind), and a call beyond the scratch's limit refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import tileink_amd as tk
from tileink_amd import _abi, scenes
from tileink_amd._abi import check, lib

pytestmark = pytest.mark.gpu

GUARD = 0xA5A5A5A5


def _frame_inputs(ctx, canvas):
    """The scene's coarse inputs on the device, the backdrops and segment ranges
    from the frame's own scan, and the frame's coarse outputs for reference."""
    scene = canvas.encode(0)
    arrays = tk.SceneArrays(scene)
    frame = tk.Frame(ctx, scene)
    frame.set_capture(True)
    frame.render()
    B, Q = scene.backdrop_len, scene.batch_count
    T = scene.tiles_width * scene.tiles_height
    K = int(frame.stats()["ptcl_records"])
    frame_out = {
        "counts": frame.buffer(_abi.BUF_TILE_PTCL_COUNTS, np.uint32, Q * T),
        "kinds": frame.buffer(_abi.BUF_TILE_KINDS, np.uint32, Q * T),
        "ptcl": frame.buffer(_abi.BUF_PTCL, np.uint32, 6 * K),
    }
    host = {
        "draws": arrays.draws, "paths": arrays.paths,
        "backdrops": frame.buffer(_abi.BUF_BACKDROPS, np.int32, B),
        "segment_ranges": frame.buffer(_abi.BUF_SEGMENT_RANGES, np.uint32, 2 * B),
        "layers": arrays.layers, "batches": arrays.batches, "paint": arrays.paint,
        "draw_batch_ids": arrays.draw_batch_ids, "tile_draws": arrays.tile_draws,
        "tile_draw_indices": arrays.tile_draw_indices,
    }
    del frame
    dev = {k: ctx.to_device(np.ascontiguousarray(v).view(np.uint32).reshape(-1) if np.asarray(v).size
                            else np.zeros(1, np.uint32)) for k, v in host.items()}
    cfg = _abi.CoarseConfig(tile_count=T, tiles_width=scene.tiles_width, tiles_height=scene.tiles_height,
                            batch_count=Q, draw_count=scene.draw_count, path_count=scene.path_count,
                            paint_words=scene.paint_words, paint_brush_base=scene.paint_brush_base,
                            layer_count=scene.layer_count, tile_draw_index_count=scene.tile_draw_index_count,
                            ptcl_capacity=scene.ptcl_capacity)
    ins = _abi.CoarseInputs(*[dev[k].ptr for k in ("draws", "paths", "backdrops", "segment_ranges", "layers",
                                                    "batches", "paint", "draw_batch_ids", "tile_draws",
                                                    "tile_draw_indices")])
    return cfg, ins, dev, frame_out, K


def _run(ctx, entry, cfg, ins):
    Q, T, cap = cfg.batch_count, cfg.tile_count, int(cfg.ptcl_capacity)
    bufs = {
        "counts": ctx.to_device(np.full(Q * T, GUARD, np.uint32)),
        "ranges": ctx.to_device(np.full(2 * Q * T, GUARD, np.uint32)),
        "kinds": ctx.to_device(np.full(Q * T, GUARD, np.uint32)),
        "bases": ctx.to_device(np.full(Q + 1, GUARD, np.uint32)),
        "ptcl": ctx.to_device(np.full(6 * cap, GUARD, np.uint32)),
        "status": ctx.to_device(np.zeros(4, np.uint32)),
    }
    outs = _abi.CoarseOutputs(*[bufs[k].ptr for k in ("counts", "ranges", "kinds", "bases", "ptcl", "status")])
    check(getattr(lib(), entry)(ctx.handle, C.byref(cfg), C.byref(ins), C.byref(outs)))
    ctx.synchronize()
    sizes = {"counts": Q * T, "ranges": 2 * Q * T, "kinds": Q * T, "bases": Q + 1, "ptcl": 6 * cap, "status": 4}
    return {k: bufs[k].read(np.uint32, n) for k, n in sizes.items()}


def _check_scene(ctx, canvas):
    cfg, ins, dev, frame_out, K = _frame_inputs(ctx, canvas)
    want = _run(ctx, "tk_coarse", cfg, ins)
    got = _run(ctx, "tk_coarse_no_alloc", cfg, ins)
    for key in want:
        assert np.array_equal(got[key], want[key]), key
    assert not got["status"][0], "overflow"
    assert int(got["status"][1]) == K == int(got["bases"][-1])
    # and both equal the frame's own coarse stage
    assert np.array_equal(got["counts"], frame_out["counts"])
    assert np.array_equal(got["kinds"], frame_out["kinds"])
    assert np.array_equal(got["ptcl"][:6 * K], frame_out["ptcl"])
    return cfg, ins, dev


def test_no_alloc_coarse_equals_tk_coarse_on_two_calls(ctx):
    big = scenes.mixed_scene_640x360()
    print("no-alloc coarse: 640x360 mixed scene", flush=True)
    cfg1, _, _ = _check_scene(ctx, big)
    small = tk.Canvas(96, 64)
    small.push_clip_sdf_rect_layer(4.5, 3.25, 90.0, 60.0, 6.0)
    small.push_path(tk.BezPath().move_to(5, 5).curve_to(90, 0, 0, 70, 80, 60).close_path(), tk.Color(0.9, 0.2, 0.1, 0.7))
    small.push_rect(10, 10, 60, 40, 4.0, tk.Color(0.1, 0.3, 0.9, 1.0))
    small.pop_layer()
    print("no-alloc coarse: 96x64 scene after it", flush=True)
    cfg2, _, _ = _check_scene(ctx, small)
    assert cfg2.tile_count * cfg2.batch_count < cfg1.tile_count * cfg1.batch_count


def test_no_alloc_coarse_refuses_tables_beyond_its_scratch(ctx):
    canvas = tk.Canvas(64, 64)
    canvas.push_rect(4, 4, 60, 60, 2.0, tk.Color(1, 0, 0, 1))
    cfg, ins, dev, _, _ = _frame_inputs(ctx, canvas)
    limit = _abi.TK_NO_ALLOC_MAX_TABLE
    cfg.tiles_width, cfg.tiles_height = 4096, 4096
    cfg.tile_count = 4096 * 4096
    cfg.batch_count = limit // cfg.tile_count + 1
    status = ctx.to_device(np.zeros(4, np.uint32))
    outs = _abi.CoarseOutputs(status.ptr, status.ptr, status.ptr, status.ptr, status.ptr, status.ptr)
    rc = lib().tk_coarse_no_alloc(ctx.handle, C.byref(cfg), C.byref(ins), C.byref(outs))
    assert rc == _abi.TK_ERR_INVALID
    assert b"exceeds" in lib().tk_last_error()
    ctx.synchronize()
    assert status.read(np.uint32, 4).tolist() == [0, 0, 0, 0]
