"""Text outlines (PTCL_PATH_GLYPH) on the gfx950 device against the oracle: every
integer stage bit-exact, and the image within 1 LSB.  The text composite raises
to the power 2.4 and 1/2.4 (fine/text/gamma.hlsli); powf is the device library's
on the GPU and libm's in the oracle, both within an ulp or two of the true power,
so a pixel whose value lies at a rounding boundary can differ by one step."""
import numpy as np
import pytest

import tileink_amd as tk
from tileink_amd import scenes

from helpers import compare_frame
from test_text import text_canvas

pytestmark = pytest.mark.gpu


def text_scene(n=300, seed=9):
    rng = np.random.default_rng(seed)
    canvas = tk.Canvas(512, 384)
    canvas.push_rect(0, 0, 512, 384, 0.0,
                     tk.LinearGradient((0, 0), (512, 384), [(0.0, tk.Color(0, 0, 0, 1)), (1.0, tk.Color(1, 1, 1, 1))]))
    for k in range(n):
        lo = rng.uniform(0, 500, 2)
        glyph = scenes._cubic_path(rng, int(rng.integers(3, 7)), tuple(lo), tuple(lo + rng.uniform(3, 16, 2)), True)
        canvas.push_text_outline(glyph, tk.Color(*rng.uniform(0, 1, 3), float(rng.uniform(0.4, 1.0))))
    return canvas


@pytest.mark.parametrize("make", [lambda: text_canvas(1.0), lambda: text_canvas(1.5), text_scene],
                         ids=["glyphs", "glyphs_scale_1.5", "many_glyphs"])
def test_text_outlines_match_oracle(ctx, make):
    canvas = make()
    gpu, ref, bad = compare_frame(ctx, canvas)
    assert [b for b in bad if b != "image"] == []
    assert (ref["ptcl"][:, 0] == 11).any()
    d = np.abs(gpu["image"].view(np.uint8).astype(np.int32) - ref["image"].view(np.uint8).astype(np.int32))
    assert d.max(initial=0) <= 1
    assert (d.reshape(-1, 4).max(1) > 0).mean() < 0.01
