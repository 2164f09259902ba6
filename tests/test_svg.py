"""SVG import (tileink_amd.svg) on the CPU: path-data parsing against SVG's
rules, and the recorded scenes rendered by the oracle against the images the
reference's own native renderer produced for its SVG test fixtures."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import svg

import svg_golden


def points(path):
    verbs, pts = path.arrays()
    return list(verbs), np.asarray(pts)


def test_packed_arc_flags_and_implicit_commands():
    # Arc flags need no separators; repeated coordinates after M are LineTo.
    a = svg.parse_path("M 100 100 h 25 a 25 25 0 1 0 -25 25 z")
    b = svg.parse_path("M100 100h25a25 25 0 10-25 25z")
    assert points(a)[0] == points(b)[0] and np.allclose(points(a)[1], points(b)[1])
    v, p = points(svg.parse_path("M 10 10 20 20 30 10"))
    assert v == [tk._abi.PATH_MOVE, tk._abi.PATH_LINE, tk._abi.PATH_LINE]


def test_parse_stops_at_the_first_error():
    v, p = points(svg.parse_path("M 30 40 L 110 160 L 150#80"))
    assert v == [tk._abi.PATH_MOVE, tk._abi.PATH_LINE] and np.allclose(p[-1], [110, 160])
    assert points(svg.parse_path("L 10 10"))[0] == []  # MoveTo is required


def test_subpath_after_close_starts_at_the_closed_start():
    v, p = points(svg.parse_path("M 10 10 L 10 20 L 20 10 z L 100 100"))
    assert v[-2:] == [tk._abi.PATH_MOVE, tk._abi.PATH_LINE] and np.allclose(p[-2], [10, 10])


def test_arc_endpoints_and_radius():
    # A half circle of radius 25 from (0, 0) to (50, 0): its points stay on the circle.
    v, p = points(svg.parse_path("M 0 0 A 25 25 0 0 1 50 0"))
    assert np.allclose(p[-1], [50, 0])
    ends = p[3::3]  # cubic end points
    assert np.allclose(np.hypot(ends[:, 0] - 25, ends[:, 1]), 25)


def test_unsupported_content_is_refused():
    with pytest.raises(svg.SvgUnsupported):
        svg.load('<svg xmlns="http://www.w3.org/2000/svg" viewBox="0 0 10 10"><text>a</text></svg>', 100)
    with pytest.raises(svg.SvgUnsupported):
        svg.load('<svg xmlns="http://www.w3.org/2000/svg" viewBox="0 0 10 10">'
                 '<path d="M0 0L5 5" stroke="red" stroke-dasharray="1 2"/></svg>', 100)


@pytest.mark.parametrize("name", svg_golden.fixtures())
def test_oracle_render_matches_reference_golden(name):
    canvas = svg.load(svg_golden.read_svg(name), 300)
    image, _ = oracle.render_frame(canvas.arrays(0))
    svg_golden.assert_close(image, name)


def test_tile_edge_area_pixel_kat():
    # native_tile_edge_area_preserves_half_channel_rounding (renderer_rounding_gpu.rs:76-88):
    # shapes/line/simple-case.svg at 300 px, pixel (103, 132) is exactly [63, 71, 0, 205].
    canvas = svg.load(svg_golden.read_svg("shapes/line/simple-case.svg"), 300)
    image, _ = oracle.render_frame(canvas.arrays(0))
    assert list(image.view(np.uint8).reshape(300, 300, 4)[132, 103]) == [63, 71, 0, 205]
