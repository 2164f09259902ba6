"""Image brushes on the CPU: the atlas the encoder packs (resource_brush_header,
src/shared/gpu_brush.rs:92-127), the image rect of Canvas::push_image
(PatternBrush::for_rect_resource, src/shared/brush.rs:557-583) and the oracle's
image sampling (sample_resource_pattern_atlas, shared/brush/pattern.hlsli:9-55),
pinned to closed forms:

- premultiplication by mul_div255 is round(c * a / 255) for every pair (exhaustive);
- nearest sampling of an image drawn at its own size copies its texels, and at
  twice its size repeats each texel over 2 x 2 pixels;
- pad / repeat / reflect pick texel clamp(i) / i mod n / the mirrored index, for
  negative indices too (euclidean remainder);
- bilinear sampling half-way between texel centres is the rounded mean of the two
  (four) neighbours: (a + b + 1) // 2 per channel, and for pad the neighbours are
  clamped to the image;
- opacity scales each premultiplied channel by mul_div255.
"""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import TileinkError


def random_image(rng, h, w, opaque=True):
    rgba = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    if opaque:
        rgba[..., 3] = 255
    return tk.Image.from_rgba8(rgba)


def render(canvas, clear=0):
    image, _ = oracle.render_frame(canvas.arrays(clear), outputs=False)
    return image


def channels(words):
    return np.asarray(words, np.uint32)[..., None].view(np.uint8).reshape(*np.shape(words), 4).astype(np.int64)


def pack(ch):
    ch = np.asarray(ch, np.uint32)
    return ch[..., 0] | (ch[..., 1] << 8) | (ch[..., 2] << 16) | (ch[..., 3] << 24)


def extend_index(i, n, extend):
    if extend == 0:
        return np.clip(i, 0, n - 1)
    if extend == 1:
        return np.mod(i, n)
    m = np.mod(i, 2 * n)
    return np.where(m < n, m, 2 * n - 1 - m)


def test_premultiply_is_rounded_product():
    c, a = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    rgba = np.stack([c, c, c, a], -1).astype(np.uint8)
    got = channels(tk.Image.from_rgba8(rgba).pixels)
    expect = np.floor(c * a / 255.0 + 0.5).astype(np.int64)  # c * a / 255 is never a half
    assert np.array_equal(got[..., 0], expect)
    assert np.array_equal(got[..., 3], a)


@pytest.mark.parametrize("scale", [1, 2])
def test_nearest_image_copies_texels(scale):
    rng = np.random.default_rng(11)
    im = random_image(rng, 12, 20)
    canvas = tk.Canvas(64, 48)
    image_id = canvas.add_image(im)
    x0, y0 = 8, 5
    assert canvas.push_image(x0, y0, x0 + 20 * scale, y0 + 12 * scale, image_id, sampling=tk.PATTERN_NEAREST) is not None
    out = render(canvas)
    region = out[y0:y0 + 12 * scale, x0:x0 + 20 * scale]
    assert np.array_equal(region, np.repeat(np.repeat(im.pixels, scale, 0), scale, 1))
    mask = np.ones(out.shape, bool)
    mask[y0:y0 + 12 * scale, x0:x0 + 20 * scale] = False
    assert not out[mask].any()


@pytest.mark.parametrize("extend", [0, 1, 2])
def test_nearest_extend_modes(extend):
    rng = np.random.default_rng(3 + extend)
    im = random_image(rng, 8, 4)
    canvas = tk.Canvas(40, 36)
    image_id = canvas.add_image(im)
    # user space -> unit square: x / 4 - 3 / 4, y / 8 - 5 / 8, so pixel centre
    # (i + 0.5, j + 0.5) samples texel (i - 3, j - 5) before the extend
    brush = tk.ImageBrush(image_id, (0.25, 0.0, 0.0, 0.125, -0.75, -0.625), extend=extend,
                          sampling=tk.PATTERN_NEAREST)
    canvas.push_rect(0, 0, 40, 36, 0.0, brush)
    out = render(canvas)
    jj, ii = np.meshgrid(np.arange(36), np.arange(40), indexing="ij")
    expect = im.pixels[extend_index(jj - 5, 8, extend), extend_index(ii - 3, 4, extend)]
    assert np.array_equal(out, expect)


@pytest.mark.parametrize("extend", [1, 2])
def test_bilinear_half_texel_is_rounded_mean(extend):
    rng = np.random.default_rng(21 + extend)
    im = random_image(rng, 8, 4, opaque=False)
    canvas = tk.Canvas(24, 32)
    image_id = canvas.add_image(im)
    # pixel centre (i + 0.5) maps to texel coordinate i, half-way between the
    # centres of texels i - 1 and i
    brush = tk.ImageBrush(image_id, (0.25, 0.0, 0.0, 0.125, -0.125, -0.0625), extend=extend,
                          sampling=tk.PATTERN_BILINEAR)
    canvas.push_rect(0, 0, 24, 32, 0.0, brush)
    out = render(canvas)
    jj, ii = np.meshgrid(np.arange(32), np.arange(24), indexing="ij")
    t = channels(im.pixels)
    x0, x1 = extend_index(ii - 1, 4, extend), extend_index(ii, 4, extend)
    y0, y1 = extend_index(jj - 1, 8, extend), extend_index(jj, 8, extend)
    top = (t[y0, x0] + t[y0, x1] + 1) // 2
    bottom = (t[y1, x0] + t[y1, x1] + 1) // 2
    sample = (top + bottom + 1) // 2
    # the sample is premultiplied; over a clear of 0 it is the pixel
    assert np.array_equal(out, pack(sample))


def test_bilinear_pad_clamps_neighbours():
    rng = np.random.default_rng(5)
    im = random_image(rng, 8, 8, opaque=False)
    canvas = tk.Canvas(40, 40)
    image_id = canvas.add_image(im)
    # texel coordinate (i + 0.5) / 4: fractions in eighths after the half-texel
    # shift, so every weight is an exact multiple of 1/8 and the filter's value
    # is the exact bilinear mean; past 8 the position clamps to the image edge
    brush = tk.ImageBrush(image_id, (1 / 32, 0.0, 0.0, 1 / 32, 0.0, 0.0), extend=0, sampling=tk.PATTERN_BILINEAR)
    canvas.push_rect(0, 0, 40, 40, 0.0, brush)
    out = channels(render(canvas))
    t = channels(im.pixels)
    expect = np.zeros((40, 40, 4), np.int64)
    for j in range(40):
        for i in range(40):
            qx, qy = min((i + 0.5) / 4, 8.0) - 0.5, min((j + 0.5) / 4, 8.0) - 0.5
            bx, by = int(np.floor(qx)), int(np.floor(qy))
            fx, fy = qx - bx, qy - by
            xs = [min(max(bx + k, 0), 7) for k in (0, 1)]
            ys = [min(max(by + k, 0), 7) for k in (0, 1)]
            v = ((1 - fx) * (1 - fy) * t[ys[0], xs[0]] + fx * (1 - fy) * t[ys[0], xs[1]] +
                 (1 - fx) * fy * t[ys[1], xs[0]] + fx * fy * t[ys[1], xs[1]])
            expect[j, i] = np.floor(v + 0.5)
    assert np.array_equal(out, expect)
    # beyond the image the edge texels alone are read
    assert np.array_equal(out[36:, 36:].reshape(-1, 4), np.repeat(t[7:, 7], 16, 0))


def test_bilinear_pad_resolves_position_to_eighth_bits():
    """The padded bilinear filter is the reference's hardware sampler: position to
    8 fractional bits, weights k / 256.  A fraction of 0.7222 texel weighs 185 / 256
    (0.72266), not 0.7222: at the reference's own pixel KAT (below, through the SVG
    importer) that is the difference between 27 and 28."""
    rng = np.random.default_rng(6)
    im = random_image(rng, 2, 2, opaque=True)
    canvas = tk.Canvas(4, 4)
    image_id = canvas.add_image(im)
    # pixel (0, 0) samples texel coordinate (0.5 + f, 0.5): weight of the right
    # texel = f resolved to 1/256; f = 0.3 + 1/1024 (between 77/256 and 78/256)
    f = 0.3 + 1 / 1024
    xf = (np.float32((0.5 + f) / 2 / 0.5), 0.0, 0.0, 0.5, 0.0, 0.0)
    canvas.push_rect(0, 0, 1, 1, 0.0, tk.ImageBrush(image_id, xf, extend=0, sampling=tk.PATTERN_BILINEAR))
    out = channels(render(canvas))[0, 0]
    f32 = np.float32
    px = f32(f32(f32(0.5) * f32(xf[0])) * f32(2))
    k = int(np.rint((float(px) - 0.5 - np.floor(float(px) - 0.5)) * 256))
    t = channels(im.pixels)
    expect = ((256 - k) * 256 * t[0, 0] + k * 256 * t[0, 1] + 32768) >> 16
    assert np.array_equal(out, expect)


def test_opacity_scales_premultiplied_channels():
    rng = np.random.default_rng(8)
    im = random_image(rng, 4, 4, opaque=False)
    canvas = tk.Canvas(16, 16)
    image_id = canvas.add_image(im)
    brush = tk.ImageBrush(image_id, (0.25, 0.0, 0.0, 0.25, 0.0, 0.0), extend=1, sampling=tk.PATTERN_NEAREST,
                          opacity=100)
    canvas.push_rect(0, 0, 16, 16, 0.0, brush)
    out = render(canvas)
    t = channels(im.pixels)
    scaled = t * 100 + 128
    scaled = (scaled + (scaled >> 8)) >> 8
    jj, ii = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    assert np.array_equal(out, pack(scaled[jj % 4, ii % 4]))


def test_image_over_background_composites():
    rng = np.random.default_rng(9)
    im = random_image(rng, 4, 4, opaque=False)
    canvas = tk.Canvas(16, 16)
    canvas.push_rect(0, 0, 16, 16, 0.0, tk.Color(0.2, 0.4, 0.6, 1.0))
    image_id = canvas.add_image(im)
    canvas.push_image(4, 4, 8, 8, image_id, sampling=tk.PATTERN_NEAREST)
    out = channels(render(canvas))
    bg = out[0, 0]
    src = channels(im.pixels)
    inv = 1.0 - src[..., 3:4] / 255.0
    # src-over in unit floats, rounded once (within 1 LSB of the f64 composite)
    expect = np.floor(np.clip((src / 255.0 + bg / 255.0 * inv), 0, 1) * 255.0 + 0.5)
    assert np.abs(out[4:8, 4:8] - expect).max() <= 1


def test_atlas_placement_patches_headers():
    rng = np.random.default_rng(2)
    ims = [random_image(rng, 5, 7), random_image(rng, 3, 11), random_image(rng, 9, 2)]
    canvas = tk.Canvas(32, 32)
    ids = [canvas.add_image(im) for im in ims]
    assert ids == [0, 1, 2]
    for k in (2, 0, 1, 2):  # image 2 drawn twice
        canvas.push_image(k, k, k + 10, k + 10, ids[k], extend=k % 3, sampling=k % 2)
    arrays = canvas.arrays()
    s = arrays.scalars
    assert (s["atlas_width"], s["atlas_height"], s["atlas_pages"]) == (11, 17, 1)
    atlas = np.asarray(arrays.atlas).reshape(17, 11)
    rows = [0, 5, 8]
    for im, y in zip(ims, rows):
        assert np.array_equal(atlas[y:y + im.height, :im.width], im.pixels)
        assert not atlas[y:y + im.height, im.width:].any()
    paint, base = np.asarray(arrays.paint, np.uint32), int(arrays.paint_brush_base)
    seen = []
    for d in np.asarray(arrays.draws):
        h = paint[base + int(d[6]):base + int(d[6]) + 9]
        assert h[0] == tk._abi.BRUSH_IMAGE
        seen.append(tuple(int(v) for v in h[2:7]))
    expect = {k: (0, rows[k], 0, ims[k].width, ims[k].height) for k in range(3)}
    assert seen == [expect[2], expect[0], expect[1], expect[2]]


def test_atlas_pages():
    rng = np.random.default_rng(4)
    tall = [random_image(rng, 3000, 2), random_image(rng, 3000, 3)]
    canvas = tk.Canvas(16, 16)
    ids = [canvas.add_image(im) for im in tall]
    canvas.push_image(0, 0, 8, 16, ids[0], sampling=tk.PATTERN_NEAREST)
    canvas.push_image(8, 0, 16, 16, ids[1], sampling=tk.PATTERN_NEAREST)
    arrays = canvas.arrays()
    s = arrays.scalars
    assert (s["atlas_width"], s["atlas_height"], s["atlas_pages"]) == (3, 4096, 2)
    atlas = np.asarray(arrays.atlas).reshape(2, 4096, 3)
    assert np.array_equal(atlas[1, :3000, :3], tall[1].pixels)
    out = render(canvas)
    # nearest at x / 8 * width: column i of the right half samples texel
    # floor((i + 0.5) / 8 * 3); rows sample floor((j + 0.5) / 16 * 3000)
    j = np.arange(16)
    rows = np.floor((j + 0.5) / 16 * 3000).astype(int)
    cols = np.floor((np.arange(8) + 0.5) / 8 * 3).astype(int)
    assert np.array_equal(out[:, 8:], tall[1].pixels[rows][:, cols])


def test_image_errors():
    canvas = tk.Canvas(16, 16)
    with pytest.raises(TileinkError):
        canvas.add_image(np.zeros((0, 3), np.uint32))
    with pytest.raises(TileinkError):
        canvas.push_image(0, 0, 4, 4, 0)  # no such image
    image_id = canvas.add_image(np.full((2, 2), 0xFF0000FF, np.uint32))
    with pytest.raises(TileinkError):
        canvas.push_image(0, 0, 4, 4, image_id, extend=3)
    assert canvas.push_image(4, 4, 4, 8, image_id) is None  # empty rect: nothing recorded
    assert canvas.push_image(0, 0, float("nan"), 8, image_id) is None
    with pytest.raises(TileinkError):
        canvas.push_rect(0, 0, 4, 4, 0.0, tk.ImageBrush(image_id + 1, (1, 0, 0, 1, 0, 0)))
    with pytest.raises(TileinkError):
        canvas.push_rect(0, 0, 4, 4, 0.0, tk.ImageBrush(image_id, (float("inf"), 0, 0, 1, 0, 0)))
