"""The reference's SVG test fixtures through the gfx950 pipeline: every fixture
tileink_amd.svg takes, rendered on the device, against the image the reference's
own native renderer produced for it (tests/golden/svg), and bit for bit against
the oracle on a sample."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import svg

import svg_golden

pytestmark = pytest.mark.gpu


def render(ctx, canvas):
    frame = tk.Frame(ctx, canvas.encode(0))
    frame.render()
    assert frame.stats()["overflow"] == 0
    return frame.image()


def test_every_fixture_matches_reference_golden(ctx):
    worst = []
    for name in svg_golden.fixtures():
        image = render(ctx, svg_golden.load_fixture(name))
        over8, over32, mean = svg_golden.compare(image, name)
        worst.append((over8, over32, mean, name))
        assert over8 <= svg_golden.MAX_SHARE_OVER_8 and over32 <= svg_golden.MAX_SHARE_OVER_32 and \
            mean <= svg_golden.MAX_MEAN, (name, over8, over32, mean)
    assert len(worst) >= 400


@pytest.mark.parametrize("name", svg_golden.fixtures()[::9])
def test_device_equals_oracle_on_fixtures(ctx, name):
    canvas = svg_golden.load_fixture(name)
    expected, _ = oracle.render_frame(canvas.arrays(0))
    assert np.array_equal(render(ctx, canvas), expected)


def test_tile_edge_area_pixel_kat(ctx):
    # native_tile_edge_area_preserves_half_channel_rounding (renderer_rounding_gpu.rs:76-88)
    image = render(ctx, svg_golden.load_fixture("shapes/line/simple-case.svg"))
    assert list(image.view(np.uint8).reshape(300, 300, 4)[132, 103]) == [63, 71, 0, 205]


def test_pattern_transform_pixel_kat(ctx):
    # native_pattern_transform_preserves_product_residual (renderer_rounding_gpu.rs:90-99)
    image = render(ctx, svg_golden.load_fixture("structure/image/with-transform.svg"))
    assert list(image.view(np.uint8).reshape(300, 300, 4)[135, 128]) == [28, 28, 28, 255]


def test_gradient_ramp_pixel_kat(ctx):
    # native_gradient_ramp_preserves_fused_fraction (renderer_rounding_gpu.rs:57-68)
    image = render(ctx, svg_golden.load_fixture("paint-servers/linearGradient/attributes-via-xlink-href.svg"))
    assert list(image.view(np.uint8).reshape(300, 300, 4)[30, 121]) == [237, 237, 237, 255]
