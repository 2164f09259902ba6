"""Shared helpers of the SVG golden tests: the fixtures under tests/golden/svg
(the reference's SVG test inputs with the images its native renderer produced
for them) and the comparison both the CPU and the GPU tests make."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "svg")

# The goldens are backend renders, and the geometry front ends differ where the
# reference hands curves and strokes to kurbo (offset curves) and this recorder
# flattens first (parity of exact vertices is not pinned): bound the share of
# pixels that differ, not their values.
MAX_SHARE_OVER_8 = 0.01   # of pixels differing by more than 8/255 in any premultiplied channel
MAX_SHARE_OVER_32 = 0.002
MAX_MEAN = 0.1            # mean absolute difference, 8-bit units


def fixtures():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["fixtures"]


def read_svg(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def load_fixture(name, target_width=300):
    """A fixture recorded on a Canvas; its relative image hrefs resolve beside it."""
    from tileink_amd import svg
    return svg.load(read_svg(name), target_width, base_dir=os.path.dirname(os.path.join(GOLDEN, name)))


def golden_premultiplied(name):
    """The golden PNG (straight RGBA8) as premultiplied RGBA8 bytes (H, W, 4)."""
    from PIL import Image
    g = np.asarray(Image.open(os.path.join(GOLDEN, name[:-4] + ".native.png")).convert("RGBA")).astype(np.int64)
    return np.concatenate([(g[..., :3] * g[..., 3:4] + 127) // 255, g[..., 3:4]], -1)


def compare(image_words, name):
    """(share > 8, share > 32, mean) of a premultiplied RGBA8 image against the golden."""
    ours = np.ascontiguousarray(image_words, np.uint32).view(np.uint8).reshape(image_words.shape + (4,))
    g = golden_premultiplied(name)
    assert ours.shape == g.shape, (name, ours.shape, g.shape)
    d = np.abs(ours.astype(np.int64) - g)
    worst = d.max(-1)
    return float((worst > 8).mean()), float((worst > 32).mean()), float(d.mean())


def assert_close(image_words, name):
    over8, over32, mean = compare(image_words, name)
    assert over8 <= MAX_SHARE_OVER_8 and over32 <= MAX_SHARE_OVER_32 and mean <= MAX_MEAN, (name, over8, over32, mean)
