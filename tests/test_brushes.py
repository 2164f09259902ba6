"""Gradient brushes on the CPU: the encoder (push_encoded_brush after physical_brush)
and the oracle's sample_brush, pinned to the reference's own expectations and to
closed forms.

- Sweep: the expected grey levels of native_sweep_* (src/native/runtime/tests/
  sweep_gpu.rs:9-74), computed there in f64 from atan2 and the extend mode, within
  1 LSB; the encoder's orientation reduction and signed span
  (src/shared/brush/tests.rs:8-45).
- Radial: t solved in f64 from |p - c(t)| = r(t) with the larger admissible root
  (the two-point conical definition), within 1 LSB; points outside the cone are
  transparent.
- Four-corner: exact corner colours and the bilinear centre.
"""
import math

import numpy as np
import pytest

import oracle
import tileink_amd as tk

BLACK, WHITE = tk.Color(0, 0, 0, 1), tk.Color(1, 1, 1, 1)
F32_MAX = float(np.finfo(np.float32).max)
F32_EPS = float(np.finfo(np.float32).eps)


STOPS = [(0.0, tk.Color(1, 0, 0, 1)), (0.3, tk.Color(0, 1, 0, 0.6)), (0.7, tk.Color(0, 0, 1, 1)),
         (1.0, tk.Color(1, 1, 0, 0.3))]
BRUSHES = {
    "solid": tk.Color(0.2, 0.4, 0.6, 0.8),
    "linear": tk.LinearGradient((3, 5), (120, 70), STOPS, 2),
    "radial": tk.RadialGradient((40, 50), 4.0, (70, 60), 55.0, STOPS, 1),
    "radial_focal": tk.RadialGradient((60, 60), 0.0, (64, 58), 40.0, STOPS, 0),
    "radial_cone": tk.RadialGradient((20, 20), 10.0, (90, 30), 10.0, STOPS, 0),
    "sweep": tk.SweepGradient((64, 64), 0.5, 5.0, STOPS, 1),
    "sweep_reversed": tk.SweepGradient((30, 90), 2.0, -1.0, STOPS, 2),
    "four_corner": tk.FourCornerGradient((10, 10, 110, 90), [c for _, c in STOPS]),
}


def encoded(brush, scale=1.0):
    """(paint words, paint_brush_base, data_base) of a brush pushed on a canvas."""
    canvas = tk.Canvas(16, 16, scale=scale)
    canvas.push_rect(0, 0, 16, 16, 0.0, brush)
    arrays = canvas.arrays()
    return np.array(arrays.paint, np.uint32), int(arrays.paint_brush_base), int(arrays.draws[0, 6])


def sample(brush, points, scale=1.0):
    paint, base, data = encoded(brush, scale)
    pts = np.asarray(points, np.float32).reshape(-1, 2)
    req = np.zeros((len(pts), 4), np.uint32)
    req[:, 0] = data
    req[:, 1:3] = pts.view(np.uint32)
    return oracle.brush_words(paint, base, req), (paint, base, req)


def params_of(brush, scale=1.0):
    paint, base, data = encoded(brush, scale)
    start = base + data
    return paint[start:start + 9], paint[start + 9:start + 21], paint[start + 21:]


def f32(v):
    with np.errstate(over="ignore"):  # f64 -> f32 overflows to +-inf, as Rust's `as f32` does
        return float(np.float32(v))


def extend_t(t, extend):
    if extend == 0:
        return min(max(t, 0.0), 1.0)
    if extend == 1:
        return t % 1.0
    phase = t % 2.0
    return min(phase, 2.0 - phase)


def grey_close(actual, level):
    b = np.array([actual], np.uint32).view(np.uint8)
    return all(abs(int(x) - y) <= 1 for x, y in zip(b, (level, level, level, 255)))


SWEEP_POINTS = [[1.0, 0.0], [0.0, 1.0], [-1.0, 0.0], [0.0, -1.0], [1.0, 1.0], [-1.0, 1.0], [-1.0, -1.0],
                [1.0, -1.0], [0.2, -0.7], [0.0, 0.0]]
SWEEP_CASES = [(0.0, f32(math.tau)), (-7.0, -3.0), (7.0, 3.0), (10.0, 10.5), (-10.0, -10.5), (10000.0, 10006.0),
               (-10000.0, -10006.0), (0.0, -f32(math.tau)), (3.0, 3.0), (0.0, F32_EPS / 2.0),
               (1.0e30, 1.1e30), (-1.0e30, -1.1e30), (-F32_MAX, F32_MAX), (F32_MAX, -F32_MAX)]


def sweep_expected(start, end, point, extend):
    """sweep_gpu.rs:36-59, in f64."""
    start, end = f32(start), f32(end)
    span = end - start
    if abs(span) <= F32_EPS:
        t = 0.0
    else:
        angle = 0.0 if point == [0.0, 0.0] else math.atan2(f32(point[1]), f32(point[0]))
        relative = angle - (start % math.tau)
        distance = relative % math.tau if span > 0.0 else -((-relative) % math.tau)
        t = distance / span
    return int(math.floor(extend_t(t, extend) * 255.0 + 0.5))


@pytest.mark.parametrize("start,end", SWEEP_CASES)
@pytest.mark.parametrize("extend", [0, 1, 2])
def test_oracle_sweep_matches_reference_expectations(start, end, extend):
    brush = tk.SweepGradient((0.0, 0.0), start, end, [(0.0, BLACK), (1.0, WHITE)], extend)
    out, _ = sample(brush, SWEEP_POINTS)
    for point, actual in zip(SWEEP_POINTS, out):
        level = sweep_expected(start, end, point, extend)
        assert grey_close(actual, level), (start, end, extend, point, hex(actual), level)


@pytest.mark.parametrize("start,end", [(0.25, 1.75), (-7.0, -3.0), (7.0, 3.0), (10000.0, 10006.0),
                                       (-10000.0, -10006.0), (1.0e30, 1.1e30), (-1.0e30, -1.1e30),
                                       (-F32_MAX, F32_MAX), (F32_MAX, -F32_MAX), (3.0, 3.0)])
def test_sweep_encoding_normalizes_orientation_and_keeps_the_signed_span(start, end):
    """src/shared/brush/tests.rs:8-45."""
    _, params, _ = params_of(tk.SweepGradient((2.0, 3.0), start, end, [(0.0, BLACK), (1.0, WHITE)]))
    p = params.view(np.float32)
    s, e = f32(start), f32(end)
    reduced = s % math.tau
    expected = f32(reduced - math.tau if reduced > math.pi else reduced)
    assert abs(float(p[2])) <= math.pi
    assert float(p[2]) == expected
    assert float(p[3]) == f32(e - s)
    assert float(p[0]) == 2.0 and float(p[1]) == 3.0


def test_sweep_rejects_nonfinite_angles():
    canvas = tk.Canvas(16, 16)
    for start, end in [(math.nan, 1.0), (0.0, math.inf), (-math.inf, 0.0)]:
        with pytest.raises(tk.TileinkError):
            canvas.push_rect(0, 0, 4, 4, 0.0, tk.SweepGradient((0, 0), start, end, [(0, BLACK), (1, WHITE)]))


def radial_t(c0, r0, c1, r1, p):
    """Largest t with |p - (c0 + t dc)| = r0 + t dr and r0 + t dr >= 0 (f64)."""
    dcx, dcy, dr = c1[0] - c0[0], c1[1] - c0[1], r1 - r0
    qx, qy = p[0] - c0[0], p[1] - c0[1]
    a = dcx * dcx + dcy * dcy - dr * dr
    b = -2.0 * (qx * dcx + qy * dcy + r0 * dr)
    c = qx * qx + qy * qy - r0 * r0
    if abs(a) < 1e-12:
        roots = [-c / b] if abs(b) > 1e-12 else []
    else:
        disc = b * b - 4 * a * c
        roots = [] if disc < 0 else [(-b - math.sqrt(disc)) / (2 * a), (-b + math.sqrt(disc)) / (2 * a)]
    valid = [t for t in roots if r0 + t * dr >= 0.0]
    return max(valid) if valid else None


RADIAL_CASES = [((0.0, 0.0), 0.0, (0.0, 0.0), 40.0), ((5.0, -3.0), 10.0, (5.0, -3.0), 50.0),
                ((0.0, 0.0), 5.0, (30.0, 10.0), 45.0), ((-10.0, 4.0), 20.0, (25.0, 4.0), 8.0)]


@pytest.mark.parametrize("c0,r0,c1,r1", RADIAL_CASES)
@pytest.mark.parametrize("extend", [0, 1, 2])
def test_oracle_radial_matches_two_point_conical_definition(c0, r0, c1, r1, extend):
    rng = np.random.default_rng(7)
    pts = rng.uniform(-60, 60, size=(200, 2)).astype(np.float32)
    brush = tk.RadialGradient(c0, r0, c1, r1, [(0.0, BLACK), (1.0, WHITE)], extend)
    out, _ = sample(brush, pts)
    checked = outside = 0
    for p, actual in zip(pts.astype(np.float64), out):
        t = radial_t(c0, r0, c1, r1, p)
        if t is None:
            assert actual == 0, (p, hex(actual))
            outside += 1
            continue
        level = int(math.floor(extend_t(t, extend) * 255.0 + 0.5))
        # Away from the repeat / reflect seams, where one f32 rounding flips the phase.
        if extend and min(abs(t - round(t)), abs(t * 0.5 - round(t * 0.5))) < 1e-4:
            continue
        assert grey_close(actual, level), (p, t, hex(actual), level)
        checked += 1
    assert checked >= 40 and checked + outside > 150


def test_oracle_radial_outside_the_cone_is_transparent():
    # Equal radii, disjoint circles: points far off the axis have no admissible t.
    brush = tk.RadialGradient((0.0, 0.0), 5.0, (40.0, 0.0), 5.0, [(0.0, BLACK), (1.0, WHITE)])
    out, _ = sample(brush, [[20.0, 30.0], [20.0, -30.0], [20.0, 0.0]])
    assert out[0] == 0 and out[1] == 0 and out[2] != 0


def test_oracle_four_corner_corners_and_centre():
    colors = [tk.Color(1, 0, 0, 1), tk.Color(0, 1, 0, 1), tk.Color(0, 0, 1, 1), tk.Color(1, 1, 1, 0.5)]
    brush = tk.FourCornerGradient((10.0, 20.0, 50.0, 60.0), colors)
    _, _, payload = params_of(brush)
    tl, tr, br, bl = (int(v) for v in payload[:4])
    assert [tl, tr, br] == [0xFF0000FF, 0xFF00FF00, 0xFFFF0000]
    out, _ = sample(brush, [[10, 20], [50, 20], [50, 60], [10, 60], [0, 0], [100, 100]])
    assert [int(v) for v in out] == [tl, tr, br, bl, tl, br]
    centre, _ = sample(brush, [[30, 40]])
    top, bottom = np.array([tl, tr], np.uint32).view(np.uint8), np.array([bl, br], np.uint32).view(np.uint8)
    expect = (top[:4].astype(float) + top[4:] + bottom[:4] + bottom[4:]) / 4.0
    assert np.all(np.abs(np.array([centre[0]], np.uint32).view(np.uint8) - expect) <= 1.0)


def test_canvas_scale_applies_to_sweep_centre_and_four_corner_bounds():
    _, p, _ = params_of(tk.SweepGradient((2.0, 3.0), 0.0, 1.0, [(0, BLACK), (1, WHITE)]), scale=2.0)
    assert list(p.view(np.float32)[:2]) == [4.0, 6.0]
    _, p, _ = params_of(tk.FourCornerGradient((1, 2, 3, 4), [BLACK] * 4), scale=2.0)
    assert list(p.view(np.float32)[:4]) == [2.0, 4.0, 6.0, 8.0]
    head, p, ramp = params_of(tk.RadialGradient((1, 1), 2, (3, 3), 4, [(0, BLACK), (1, WHITE)]), scale=2.0)
    assert list(p.view(np.float32)[6:12]) == [0.5, 0.0, 0.0, 0.5, 0.0, 0.0]
    assert int(head[0]) == 3 and int(head[2]) == 21 and int(head[3]) == len(ramp) >= 64


@pytest.mark.parametrize("name", sorted(BRUSHES))
def test_oracle_frame_samples_the_brush_at_pixel_centres(name):
    """Inside an axis-aligned rectangle over a transparent clear, fine writes the
    brush sample at the pixel centre (x + 0.5, y + 0.5) unchanged."""
    canvas = tk.Canvas(64, 48)
    canvas.push_rect(0, 0, 64, 48, 0.0, BRUSHES[name])
    arrays = canvas.arrays()
    image, _ = oracle.render_frame(arrays)
    ys, xs = np.mgrid[0:48, 0:64]
    pts = np.stack([xs + 0.5, ys + 0.5], -1).reshape(-1, 2)
    paint, base = np.array(arrays.paint, np.uint32), int(arrays.paint_brush_base)
    req = np.zeros((len(pts), 4), np.uint32)
    req[:, 0] = int(arrays.draws[0, 6])
    req[:, 1:3] = pts.astype(np.float32).view(np.uint32)
    expected = oracle.brush_words(paint, base, req).reshape(48, 64)
    assert np.array_equal(image, expected)


def test_gradient_interpolation_space_and_transform():
    stops = [(0.0, tk.Color(1, 0, 0, 1)), (1.0, tk.Color(0, 0, 1, 0))]
    xf = (0.5, 0.25, -0.125, 2.0, 3.0, -4.0)
    for unpremul in (False, True):
        brush = tk.LinearGradient((0, 0), (10, 0), stops, 0, transform=xf, unpremultiplied=unpremul)
        header, params, ramp = params_of(brush, scale=2.0)
        # the brush transform scaled by 1 / scale in its linear part (scale_brush_transform)
        got = params[4:10].view(np.float32)
        assert np.array_equal(got, np.float32([0.25, 0.125, -0.0625, 1.0, 3.0, -4.0]))
        n = len(ramp)
        t = np.arange(n, dtype=np.float32) / np.float32(n - 1)
        if unpremul:  # straight colour (1 - t, 0, t) with alpha 1 - t, premultiplied after
            expect = np.stack([(1 - t) * (1 - t), 0 * t, t * (1 - t), 1 - t], 1)
        else:  # premultiplied endpoints (1, 0, 0, 1) and (0, 0, 0, 0)
            expect = np.stack([1 - t, 0 * t, 0 * t, 1 - t], 1)
        words = ramp.view(np.uint8).reshape(-1, 4).astype(np.int64)
        assert np.abs(words - np.floor(expect * 255 + 0.5)).max() <= 1
