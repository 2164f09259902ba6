"""Gradient brushes on the gfx950 device: sample_brush through tk_brush_words
(validation/brush.hlsl:14-18) against the oracle and against the reference's own
sweep expectations (src/native/runtime/tests/sweep_gpu.rs:9-74), and whole frames
filled with radial, sweep and four-corner brushes against the oracle.

Bit-exact for solid, linear, radial and four-corner brushes.  The sweep's atan2f
is the device library's on the GPU and libm's in the oracle; both are within a few
ulp of the true angle, so sweep samples agree within 1 LSB per channel (the same
bound the reference's own sweep test allows)."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd._abi import check, lib

from helpers import channel_diff, compare_frame
from test_brushes import (BLACK, BRUSHES, SWEEP_CASES, SWEEP_POINTS, WHITE, encoded, grey_close,
                          sweep_expected)

pytestmark = pytest.mark.gpu


def device_brush_words(ctx, paint, base, req):
    d_paint, d_req = ctx.to_device(paint), ctx.to_device(req)
    d_out = ctx.to_device(np.full(len(req), 0xA5A5A5A5, np.uint32))
    check(lib().tk_brush_words(ctx.handle, d_paint.ptr, len(paint), base, len(req), d_req.ptr, d_out.ptr))
    ctx.synchronize()
    return d_out.read(np.uint32, len(req))


def requests(data_base, points):
    pts = np.asarray(points, np.float32).reshape(-1, 2)
    req = np.zeros((len(pts), 4), np.uint32)
    req[:, 0] = data_base
    req[:, 1:3] = pts.view(np.uint32)
    return req


@pytest.mark.parametrize("name", sorted(BRUSHES))
def test_brush_words_match_oracle(ctx, name):
    paint, base, data = encoded(BRUSHES[name])
    rng = np.random.default_rng(5)
    pts = np.concatenate([rng.uniform(-20, 150, size=(20000, 2)),
                          np.stack(np.meshgrid(np.arange(128) + 0.5, np.arange(128) + 0.5), -1).reshape(-1, 2)])
    req = requests(data, pts)
    got = device_brush_words(ctx, paint, base, req)
    ref = oracle.brush_words(paint, base, req)
    if name.startswith("sweep"):
        diff = np.abs(got.view(np.uint8).astype(int) - ref.view(np.uint8).astype(int))
        assert diff.max() <= 1, name
        assert (got != ref).mean() < 0.01, (name, (got != ref).mean())
    else:
        assert np.array_equal(got, ref), (name, int((got != ref).sum()))


def test_brush_words_out_of_range_reads_are_zero(ctx):
    paint, base, data = encoded(BRUSHES["radial"])
    # A data base past the blob reads zero words: kind 0, transparent colour.
    req = requests(len(paint) + 7, [[1.0, 2.0], [50.0, 60.0]])
    got = device_brush_words(ctx, paint, base, req)
    assert np.array_equal(got, oracle.brush_words(paint, base, req))
    assert got.tolist() == [0, 0]


@pytest.mark.parametrize("start,end", SWEEP_CASES)
def test_sweep_on_device_matches_reference_expectations(ctx, start, end):
    for extend in (0, 1, 2):
        brush = tk.SweepGradient((0.0, 0.0), start, end, [(0.0, BLACK), (1.0, WHITE)], extend)
        paint, base, data = encoded(brush)
        got = device_brush_words(ctx, paint, base, requests(data, SWEEP_POINTS))
        for point, actual in zip(SWEEP_POINTS, got):
            level = sweep_expected(start, end, point, extend)
            assert grey_close(actual, level), (start, end, extend, point, hex(actual), level)


def gradient_canvas(kinds, scale=1.0):
    canvas = tk.Canvas(160, 144, scale=scale)
    for k, name in enumerate(kinds):
        brush = BRUSHES[name]
        path = tk.BezPath().move_to(5 + 7 * k, 4).curve_to(150, 10 + 5 * k, 20, 130, 150 - 9 * k, 140).close_path()
        canvas.push_path(path, brush, rule=tk.FillRule(k % 2))
        canvas.push_rect(10 + 11 * k, 12 + 6 * k, 90 + 5 * k, 100 + 3 * k, 6.0, brush)
    canvas.push_clip_sdf_rect_layer(20.5, 15.25, 140.0, 120.0, 9.0)
    for k, name in enumerate(reversed(kinds)):
        canvas.push_stroke(tk.BezPath().move_to(10, 20 + 9 * k).line_to(150, 90 - 5 * k), 3.0 + k, BRUSHES[name])
    canvas.pop_layer()
    return canvas


@pytest.mark.parametrize("scale", [1.0, 2.0])
def test_frames_with_exact_gradients_match_oracle(ctx, scale):
    canvas = gradient_canvas(["linear", "radial", "radial_focal", "radial_cone", "four_corner", "solid"], scale)
    gpu, ref, bad = compare_frame(ctx, canvas)
    assert not bad, (bad, channel_diff(gpu["image"], ref["image"]))
    assert len(np.unique(gpu["image"])) > 200  # the gradients are sampled, not flat


def test_frames_with_sweep_gradients_match_oracle_within_one_lsb(ctx):
    canvas = gradient_canvas(["sweep", "sweep_reversed", "linear"])
    gpu, ref, bad = compare_frame(ctx, canvas)
    assert set(bad) <= {"image"}, bad
    assert channel_diff(gpu["image"], ref["image"]) <= 1
