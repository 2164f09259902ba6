"""Text outlines (path glyphs) on the CPU: the recorder's DRAW_PATH_GLYPH draws
(Canvas::push_text_layout_as_path, src/canvas/primitives/text.rs:104-146), their
PTCL_PATH_GLYPH records, and the oracle's text compositing
(fine/text/gamma.hlsli, auto.hlsli, basic.hlsli:101-115) pinned to the
mathematics:

- the sRGB transfer functions against the IEC 61966-2-1 formulas in f64, at the
  textbook values, and as inverses of each other;
- the corrected coverage in closed form for white text on black (the coverage
  gamma-encoded) and black text on white (1 - encode((1 - c)^2)), and the
  near-equal-colour branch (c + c (1 - c) decode(1 - f));
- the composite: untouched at coverage 0 or a transparent source, plain
  src-over at full coverage (exact for an opaque source), the grey reduction of
  a non-opaque destination.
"""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import scenes

F = np.float32


def bits(v):
    return np.asarray(v, np.float32).view(np.uint32)


def words(dst, src, cov, fg):
    req = np.zeros((len(dst), 4), np.uint32)
    req[:, 0], req[:, 1], req[:, 2], req[:, 3] = dst, src, cov, bits(fg)
    return oracle.text_words(req)


def decode64(v):
    v = np.clip(v, 0, 1)
    return np.where(v > 0.04045, ((v + 0.055) / 1.055) ** 2.4, v / 12.92)


def encode64(v):
    v = np.clip(v, 0, 1)
    return np.where(v > 0.0031308, 1.055 * v ** (1 / 2.4) - 0.055, v * 12.92)


def as_f32(w):
    return np.asarray(w, np.uint32).view(np.float32).astype(np.float64)


def test_srgb_transfer_functions():
    v = np.concatenate([np.linspace(0, 1, 4001), [0.04045, 0.0031308, 0.5, -0.2, 1.3]]).astype(np.float32)
    out = words(np.zeros(len(v)), np.zeros(len(v)), np.zeros(len(v)), v)
    dec, enc = as_f32(out[:, 2]), as_f32(out[:, 3])
    assert np.allclose(dec, decode64(v.astype(np.float64)), rtol=2e-6, atol=1e-7)
    assert np.allclose(enc, encode64(v.astype(np.float64)), rtol=2e-6, atol=1e-7)
    assert abs(dec[np.where(v == F(0.5))[0][0]] - 0.2140411) < 1e-6
    assert abs(enc[np.where(v == F(0.5))[0][0]] - 0.7353570) < 1e-6
    # the two are inverses
    mid = v[(v > 0) & (v < 1)]
    back = words(np.zeros(len(mid)), np.zeros(len(mid)), np.zeros(len(mid)), as_f32(words(
        np.zeros(len(mid)), np.zeros(len(mid)), np.zeros(len(mid)), mid)[:, 2]).astype(np.float32))
    assert np.allclose(as_f32(back[:, 3]), mid, atol=2e-6)


def test_corrected_coverage_closed_forms():
    c = np.arange(256)
    raw = c / 255.0
    white = as_f32(words(np.zeros(256), np.zeros(256), c, np.ones(256))[:, 1])
    black = as_f32(words(np.zeros(256), np.zeros(256), c, np.zeros(256))[:, 1])
    assert white[0] == 0 and white[255] == 1 and black[0] == 0 and black[255] == 1
    assert np.allclose(white[1:255], encode64(raw[1:255]), atol=3e-6)  # white on black: c gamma-encoded
    assert np.allclose(black[1:255], 1 - encode64((1 - raw[1:255]) ** 2), atol=3e-6)
    # foreground = background = 0.5: the inverse blend is ill-conditioned, so the
    # enhanced coverage is returned
    half = as_f32(words(np.zeros(256), np.zeros(256), c, np.full(256, 0.5))[:, 1])
    enhanced = raw + raw * (1 - raw) * decode64(0.5)
    assert np.allclose(half[1:255], enhanced[1:255], atol=2e-6)
    # the general case against the formula in f64
    rng = np.random.default_rng(1)
    fg = rng.uniform(0, 1, 4000).astype(np.float32)
    cov = rng.integers(1, 255, 4000)
    got = as_f32(words(np.zeros(4000), np.zeros(4000), cov, fg)[:, 1])
    f, r = fg.astype(np.float64), cov / 255.0
    bg = 1 - f
    e = r + r * (1 - r) * decode64(bg)
    general = np.clip((encode64(decode64(bg) + (decode64(f) - decode64(bg)) * e) - bg) / (f - bg), 0, 1)
    expect = np.where(np.abs(f - bg) < 1 / 256, e, general)
    far = np.abs(f - bg) > 0.02  # away from the branch, where f32 rounding of f - bg is amplified
    assert np.allclose(got[far], expect[far], atol=2e-4)


def pack(r, g, b, a):
    return (np.asarray(r, np.uint32) | (np.asarray(g, np.uint32) << 8) | (np.asarray(b, np.uint32) << 16) |
            (np.asarray(a, np.uint32) << 24))


def test_composite_edges():
    rng = np.random.default_rng(2)
    n = 5000
    a = rng.integers(0, 256, n)
    src = pack(*(rng.integers(0, 256, n) * a // 255 for _ in range(3)), a)
    da = rng.integers(0, 256, n)
    dst = pack(*(rng.integers(0, 256, n) * da // 255 for _ in range(3)), da)
    fg = np.zeros(n, np.float32)
    # coverage 0 or a transparent source: the destination unchanged
    assert np.array_equal(words(dst, src, np.zeros(n), fg)[:, 0], dst)
    assert np.array_equal(words(dst, src & 0x00FFFFFF, np.full(n, 255), fg)[:, 0], dst)
    # full coverage: src + dst (1 - alpha) per channel, rounded; exact for an opaque source
    out = words(dst, src, np.full(n, 255), fg)[:, 0]
    s = src[:, None] >> np.array([0, 8, 16, 24]) & 255
    d = dst[:, None] >> np.array([0, 8, 16, 24]) & 255
    al = (src >> 24)[:, None] / 255.0
    expect = np.floor(np.concatenate([s[:, :3] + d[:, :3] * (1 - al), al * 255 + d[:, 3:] * (1 - al)], 1) + 0.5)
    got = out[:, None] >> np.array([0, 8, 16, 24]) & 255
    assert np.abs(got - expect).max() <= 1
    opaque = (src >> 24) == 255
    assert np.array_equal(out[opaque], src[opaque])
    # partial coverage: the corrected coverage c in src * c + dst * (1 - alpha c),
    # alpha out = 255 alpha c + dst alpha (1 - alpha c), each rounded once
    cov = rng.integers(1, 255, n)
    # the foreground the composite takes: the source's straight colour, Rec. 709 weights
    alpha32 = np.maximum((src >> 24).astype(np.float32), F(1))
    straight = [np.clip(((src >> (8 * k)) & 255).astype(np.float32) / alpha32, 0, 1) for k in range(3)]
    fgv = (straight[0] * F(0.2126) + straight[1] * F(0.7152)) + straight[2] * F(0.0722)
    w = words(dst, src, cov, fgv)
    c = as_f32(w[:, 1])[:, None]
    amount = al * c
    expect = np.floor(np.concatenate([s[:, :3] * c + d[:, :3] * (1 - amount), amount * 255 + d[:, 3:] * (1 - amount)],
                                     1) + 0.5)
    got = w[:, 0][:, None] >> np.array([0, 8, 16, 24]) & 255
    keep = (src >> 24) != 0
    assert np.abs(got - expect)[keep].max() <= 1


def text_canvas(scale=1.0):
    rng = np.random.default_rng(5)
    canvas = tk.Canvas(96, 72, scale=scale)
    canvas.push_rect(0, 0, 48, 72, 0.0, tk.Color(0.05, 0.05, 0.1, 1.0))
    canvas.push_rect(48, 0, 96, 72, 0.0, tk.Color(0.95, 0.9, 0.85, 0.7))
    for k in range(10):
        # small closed contours, a few pixels across, as glyph outlines are
        lo = rng.uniform(0, 80, 2)
        glyph = scenes._cubic_path(rng, 4, tuple(lo), tuple(lo + rng.uniform(4, 14, 2)), True)
        brush = [tk.Color(1, 1, 1, 1), tk.Color(0, 0, 0, 1), tk.Color(0.9, 0.2, 0.1, 0.8),
                 tk.LinearGradient((0, 0), (96, 72), [(0.0, tk.Color(0, 0.4, 1, 1)), (1.0, tk.Color(1, 1, 0, 1))])][k % 4]
        canvas.push_text_outline(glyph, brush)
    canvas.push_clip_layer(scenes._cubic_path(rng, 5, (10, 10), (86, 62), True))
    canvas.push_text_outline(scenes._cubic_path(rng, 6, (20, 20), (70, 50), True), tk.Color(0.2, 0.8, 0.3, 0.9))
    canvas.pop_layer()
    return canvas


def test_text_outline_records_path_glyphs():
    canvas = text_canvas()
    arrays = canvas.arrays()
    tags = np.asarray(arrays.draws)[:, 8]
    assert (tags == 5).sum() == 11  # DRAW_PATH_GLYPH
    image, ref = oracle.render_frame(arrays)
    assert (ref["ptcl"][:, 0] == 11).any()  # PTCL_PATH_GLYPH
    assert canvas.push_text_outline(tk.BezPath(), tk.Color(1, 1, 1)) is None


def test_full_coverage_text_is_plain_source_over():
    # a square outline covering whole pixels: coverage 255, so the text composite is
    # src-over in 8 bits; an opaque brush replaces the pixel
    canvas = tk.Canvas(32, 32)
    canvas.push_rect(0, 0, 32, 32, 0.0, tk.Color(0.2, 0.4, 0.6, 1.0))
    square = tk.BezPath().move_to(8, 8).line_to(24, 8).line_to(24, 24).line_to(8, 24).close_path()
    canvas.push_text_outline(square, tk.Color(1.0, 0.0, 0.5, 1.0))
    image, _ = oracle.render_frame(canvas.arrays(), outputs=False)
    inner = image[8:24, 8:24]
    assert np.all(inner == pack(255, 0, 128, 255))
