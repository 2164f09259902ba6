"""Image brushes on the gfx950 device against the oracle, bit-exact at every
stage: image rects (PTCL_IMAGE for tiles an opaque image covers, SDF records
elsewhere), image brushes on rounded rects, fills and strokes under rotations,
every extend and sampling mode, opacity below 255, a clip layer around an
image, a canvas scale of 1.5, a 1024 x 1024 frame of a large bilinear image,
atlases merged across the scenes of a batch, and a frame updated with new
image contents."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import scenes

from helpers import compare_frame

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,seed", [(1.0, 0), (1.5, 1), (1.0, 2)])
def test_image_scene_matches_oracle(ctx, scale, seed):
    canvas = scenes.image_scene(scale=scale, seed=seed)
    gpu, ref, bad = compare_frame(ctx, canvas)
    assert bad == []
    tags = set(np.unique(ref["ptcl"][:, 0]).tolist())
    assert {9, 13} <= tags  # image brushes as SDF records and as whole-tile images


@pytest.mark.parametrize("extend", [0, 1, 2])
@pytest.mark.parametrize("sampling", [tk.PATTERN_NEAREST, tk.PATTERN_BILINEAR])
def test_extend_and_sampling_modes(ctx, extend, sampling):
    rng = np.random.default_rng(40 + 3 * extend + sampling)
    canvas = tk.Canvas(96, 80)
    image_id = canvas.add_image(scenes.synthetic_image(rng, 13, 9))
    # rotated, scaled and shifted so texel coordinates run negative and past the image
    xf = (0.07, 0.03, -0.025, 0.09, -0.8, -0.6)
    canvas.push_rect(0, 0, 96, 80, 0.0, tk.ImageBrush(image_id, xf, extend=extend, sampling=sampling))
    canvas.push_rect(10.5, 12.25, 70.75, 60.5, 9.0,
                     tk.ImageBrush(image_id, xf, extend=extend, sampling=sampling, opacity=140))
    gpu, ref, bad = compare_frame(ctx, canvas)
    assert bad == []


def test_large_bilinear_image(ctx):
    rng = np.random.default_rng(7)
    canvas = tk.Canvas(1024, 1024)
    photo = canvas.add_image(scenes.synthetic_image(rng, 480, 640, opaque=True))
    canvas.push_image(12.3, 40.7, 1012.1, 790.4, photo)
    canvas.push_image(100.0, 800.0, 420.0, 1020.0, photo, sampling=tk.PATTERN_NEAREST)
    gpu, ref, bad = compare_frame(ctx, canvas)
    assert bad == []


def test_batch_merges_atlases(ctx):
    canvases = [scenes.image_scene(200, 150, seed=3), scenes.mixed_scene_640x360(),
                scenes.image_scene(128, 300, scale=1.25, seed=4)]
    rng = np.random.default_rng(1)
    tall = tk.Canvas(64, 64)
    ids = [tall.add_image(scenes.synthetic_image(rng, 2500, 3)) for _ in range(2)]
    tall.push_image(0, 0, 32, 64, ids[0], sampling=tk.PATTERN_NEAREST)
    tall.push_image(32, 0, 64, 64, ids[1])
    canvases.append(tall)  # two atlas pages
    encoded = [c.encode() for c in canvases]
    assert encoded[-1].atlas_pages == 2
    batch = tk.Batch(ctx, encoded)
    batch.render()
    assert batch.frame.stats()["overflow"] == 0
    for i, canvas in enumerate(canvases):
        expected, _ = oracle.render_frame(canvas.arrays(), outputs=False)
        got = batch.image(i)
        assert np.array_equal(got, expected), (i, int((got != expected).sum()))


def test_update_with_new_image_contents(ctx):
    def canvas_with(seed):
        rng = np.random.default_rng(seed)
        canvas = tk.Canvas(64, 48)
        image_id = canvas.add_image(scenes.synthetic_image(rng, 10, 12))
        canvas.push_image(4, 4, 60, 44, image_id)
        return canvas

    first, second = canvas_with(1), canvas_with(2)
    frame = tk.Frame(ctx, first.encode())
    frame.render()
    assert np.array_equal(frame.image(), oracle.render_frame(first.arrays(), outputs=False)[0])
    frame.update(second.encode())
    frame.render()
    expected = oracle.render_frame(second.arrays(), outputs=False)[0]
    assert np.array_equal(frame.image(), expected)


def test_render_to_image_with_images(ctx):
    canvas = scenes.image_scene(160, 120, seed=5)
    renderer = tk.NativeRenderer(tk.NativeBackend.Hip, 160, 120, context=ctx)
    got = renderer.render_to_image(canvas).readback()
    assert np.array_equal(got, oracle.render_frame(canvas.arrays(), outputs=False)[0])
