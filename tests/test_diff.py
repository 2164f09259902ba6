"""The differentiable form on the CPU: the FP64 oracle's areas against the exact
polygon area of another oracle routine, its analytic gradient against FP64
central finite differences (through the torch geometry: control points, stroke
widths, colours), and the geometry's recorded lines against its polygon edges."""
import math

import numpy as np
import pytest
import torch

import oracle
import tileink_amd as tk
from tileink_amd import diff


def weights(h, w, seed=0):
    return np.random.default_rng(seed).normal(size=(h, w, 4))


def clip_masks(sc, verts, counts):
    """Per painter-order draw, the product of the coverages of the clip layers
    around it (None outside every clip): an SDF layer's FP32 coverage from the
    oracle's SDF evaluation of the encoded clip draw, a path layer's FP64 area
    coverage of its FP32-rounded polygon under its fill rule; None for a scene
    without clips."""
    if not sc.clips:
        return None
    e = diff.encode(sc, verts, counts)
    arrays = tk.SceneArrays(e.scene)
    cov = [oracle.sdf_draw_coverage(arrays, cid).astype(np.float64) if len(clip) == 7 else
           oracle.path_clip_coverage(np.asarray(clip[0], np.float32), clip[1], sc.width, sc.height)
           for cid, clip in zip(e.clip_ids, sc.clips)]
    masks = []
    for d in range(len(counts)):
        m = None
        for c, clip in enumerate(sc.clips):
            if clip[-2] <= d < clip[-1]:
                m = cov[c] if m is None else m * cov[c]
        masks.append(m)
    return masks


def gradient_weights(line, offsets, width, height, kind="linear"):
    """Each stop's weight (H, W, n) in a gradient paint at the pixel centres: t =
    the projection on the line (linear) or the distance from its start over its
    length (radial), clamped to [0, 1], or the angle about its start from its
    direction over a full turn (sweep); linear between stops, constant beyond the
    end stops.  A four-corner paint is bilinear over the box the line spans."""
    x0, y0, x1, y1 = line
    X, Y = np.meshgrid(np.arange(width) + 0.5, np.arange(height) + 0.5)
    dx, dy = x1 - x0, y1 - y0
    den = dx * dx + dy * dy
    if kind == "four_corner":
        u = np.clip((X - x0) / dx, 0.0, 1.0) if dx else np.zeros_like(X)
        v = np.clip((Y - y0) / dy, 0.0, 1.0) if dy else np.zeros_like(X)
        return np.stack([(1 - u) * (1 - v), u * (1 - v), (1 - u) * v, u * v], -1)
    if kind == "sweep":
        start = math.atan2(dy, dx) if den > 0 else 0.0
        t = np.mod((np.arctan2(Y - y0, X - x0) - start) / (2 * math.pi), 1.0)
    elif den <= 0:
        t = np.zeros_like(X)
    elif kind == "radial":
        t = np.clip(np.hypot(X - x0, Y - y0) / math.sqrt(den), 0.0, 1.0)
    else:
        t = np.clip(((X - x0) * dx + (Y - y0) * dy) / den, 0.0, 1.0)
    o = list(offsets)
    n = len(o)
    w = np.zeros(t.shape + (n,))
    done = np.zeros(t.shape, bool)
    sel = (t <= o[0]) | (n == 1)
    w[sel, 0] = 1.0
    done |= sel
    for j in range(n - 1):
        sel = ~done & (t < o[j + 1])
        u = (t - o[j]) / (o[j + 1] - o[j]) if o[j + 1] > o[j] else np.zeros_like(t)
        w[sel, j] = 1.0 - u[sel]
        w[sel, j + 1] = u[sel]
        done |= sel
    w[~done, n - 1] = 1.0
    return w


def paint_colors(sc):
    """Per painter-order draw, its premultiplied colour (4,) or, for a gradient
    paint, its colour image (H, W, 4) at the pixel centres, from the current values."""
    with torch.no_grad():
        pm = sc.premultiplied().detach().cpu().double().numpy()
        out = list(pm)
        for d, entry in sc.gradients.items():
            line, offsets, stops, kind = diff.gradient_parts(entry)
            st = stops.detach().cpu().double().numpy()
            spm = np.concatenate([st[:, :3] * st[:, 3:4], st[:, 3:4]], 1)
            out[d] = gradient_weights(line, offsets, sc.width, sc.height, kind) @ spm
    return out


def analytic_grads(sc, w, recorded=False):
    """dL/d(parameters) for L = sum(w * image) from the FP64 oracle's line and colour
    gradients, carried through the geometry by torch; recorded=True evaluates them at
    the vertices rounded to FP32, where the device renders."""
    st = sc.structure()
    params = [p for p in (sc.fills, sc.strokes, sc.widths, sc.colors) if p is not None]
    params += [g[2] for g in sc.gradients.values()]
    for p in params:
        p.grad = None
        p.requires_grad_(True)
    v, k = sc.polygons(st)
    pm = sc.premultiplied()
    draws, lines, la, lb = diff.oracle_inputs(sc, v, k, recorded=recorded)
    lg, maps = oracle.diff_backward(sc.width, sc.height, diff.clear_color(sc), draws, paint_colors(sc), lines, w,
                                    masks=clip_masks(sc, v, k), color_maps=True)
    gv = torch.zeros_like(v)
    lgt = torch.as_tensor(lg)
    gv.index_add_(0, torch.as_tensor(la), lgt[:, :2])
    gv.index_add_(0, torch.as_tensor(lb), lgt[:, 2:])
    cg = np.stack([m.sum((0, 1)) for m in maps]) if maps else np.zeros((0, 4))
    tensors, grads = [v, pm], [gv, None]
    for d, entry in sc.gradients.items():
        line, offsets, stops, kind = diff.gradient_parts(entry)
        cg[d] = 0.0  # the draw's solid colour is unused
        wts = gradient_weights(line, offsets, sc.width, sc.height, kind)
        spm = torch.cat([stops[:, :3] * stops[:, 3:4], stops[:, 3:4]], 1)
        tensors.append(spm)
        grads.append(torch.as_tensor(np.einsum("hwn,hwc->nc", wts, maps[d]), dtype=spm.dtype))
    grads[1] = torch.as_tensor(cg, dtype=pm.dtype)
    torch.autograd.backward(tensors, grads)
    return st


def fd_loss(sc, st, w, masks=None):
    with torch.no_grad():
        v, k = sc.polygons(st)
        draws, lines, _, _ = diff.oracle_inputs(sc, v, k)
        image = oracle.diff_render(sc.width, sc.height, diff.clear_color(sc), draws, paint_colors(sc), lines,
                                   masks=masks)
    return float((image * w).sum())


def check_fd(sc, w, per_param=12, h=1e-3, seed=1, rtol=2e-5):
    st = analytic_grads(sc, w)
    with torch.no_grad():
        masks = clip_masks(sc, *sc.polygons(st))
    rng = np.random.default_rng(seed)
    checked = 0
    for p in (sc.fills, sc.strokes, sc.widths, sc.colors, *[g[2] for g in sc.gradients.values()]):
        if p is None:
            continue
        flat = p.data.view(-1)
        g = p.grad.view(-1)
        scale = float(g.abs().max())
        for i in rng.choice(flat.numel(), size=min(per_param, flat.numel()), replace=False):
            o = float(flat[i])
            flat[i] = o + h
            lp = fd_loss(sc, st, w, masks)
            flat[i] = o - h
            lm = fd_loss(sc, st, w, masks)
            flat[i] = o
            fd = (lp - lm) / (2 * h)
            assert abs(fd - float(g[i])) <= rtol * max(abs(fd), 0.01 * scale) + 1e-9, (i, fd, float(g[i]))
            checked += 1
    return checked


def test_oracle_area_matches_exact_polygon_area():
    rng = np.random.default_rng(1)
    for n in (3, 7, 12):
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = rng.uniform(4, 14, n)
        pts = np.stack([20.3 + rad * np.cos(ang), 17.7 + rad * np.sin(ang)], 1).astype(np.float32)
        lines = np.concatenate([pts, np.roll(pts, -1, 0)], 1).astype(np.float64)
        box, area = oracle.diff_area(lines, 0, n, 40, 36)
        rec = np.zeros((n, 6), np.float32)
        rec[:, 2:6] = lines
        words = rec.view(np.uint32).copy()
        for y in range(box[3]):
            for x in range(box[2]):
                exact = oracle.lib().or_polygon_pixel_area_f64(words.ctypes.data, n, float(box[0] + x),
                                                               float(box[1] + y))
                assert abs(exact - area[y, x]) < 1e-9
        # the areas of a simple polygon sum to its shoelace area
        shoelace = 0.5 * abs(np.sum(pts[:, 0].astype(np.float64) * np.roll(pts[:, 1], -1)
                                    - np.roll(pts[:, 0], -1).astype(np.float64) * pts[:, 1]))
        assert abs(abs(area.sum()) - shoelace) < 1e-9


def test_oracle_area_clips_to_the_canvas_and_counts_winding():
    # A square traced twice has winding 2 inside; half of it lies left of the canvas.
    sq = np.array([[-3.5, 2.25], [4.5, 2.25], [4.5, 9.75], [-3.5, 9.75]])
    ring = np.concatenate([sq, np.roll(sq, -1, 0)], 1)
    lines = np.concatenate([ring, ring])
    box, area = oracle.diff_area(lines, 0, 8, 16, 16)
    assert list(box) == [0, 2, 5, 8]
    assert np.isclose(area[1:7, 0:4], 2.0).all() or np.isclose(area[1:7, 0:4], -2.0).all()
    assert np.isclose(abs(area[0, 0]), 1.5) and np.isclose(abs(area[1, 4]), 1.0)


def test_oracle_gradient_of_one_polygon_matches_finite_differences():
    rng = np.random.default_rng(2)
    n, W, H = 9, 40, 36
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = rng.uniform(5, 14, n)
    pts = np.stack([20.3 + rad * np.cos(ang), 17.7 + rad * np.sin(ang)], 1)
    draws, cols, clear = [(0, n, 0)], np.array([[0.3, 0.2, 0.1, 0.6]]), [0.1, 0.0, 0.2, 0.3]
    w = weights(H, W, 3)
    lines = lambda p: np.concatenate([p, np.roll(p, -1, 0)], 1)
    lg, cg = oracle.diff_backward(W, H, clear, draws, cols, lines(pts), w)
    loss = lambda p, c=cols: float((oracle.diff_render(W, H, clear, draws, c, lines(p)) * w).sum())
    h = 1e-4
    for i in range(n):
        for k in range(2):
            pp, pm = pts.copy(), pts.copy()
            pp[i, k] += h
            pm[i, k] -= h
            fd = (loss(pp) - loss(pm)) / (2 * h)
            assert abs(fd - (lg[i, k] + lg[i - 1, 2 + k])) < 1e-6 * max(1.0, abs(fd))
    for k in range(4):
        cp, cm = cols.copy(), cols.copy()
        cp[0, k] += h
        cm[0, k] -= h
        assert abs((loss(pts, cp) - loss(pts, cm)) / (2 * h) - cg[0, k]) < 1e-6 * max(1.0, abs(cg[0, k]))


@pytest.mark.parametrize("seed", [3, 4])
def test_gradient_through_geometry_matches_finite_differences(seed):
    sc = diff.random_scene(seed, 96, 80, fills=5, strokes=3, cubics=4, radius_mu=math.log(14.0))
    assert check_fd(sc, weights(80, 96, seed)) >= 30


def test_opaque_overlaps_and_even_odd_match_finite_differences():
    # Opaque draws make zero factors in the transmittance; even-odd folds the winding.
    sc = diff.random_scene(7, 64, 64, fills=6, strokes=2, cubics=3, radius_mu=math.log(16.0))
    with torch.no_grad():
        sc.colors[:3, 3] = 1.0
    sc.rules[:] = 1
    sc.clear = (0.2, 0.1, 0.0, 0.4)
    assert check_fd(sc, weights(64, 64, 8)) >= 30


def test_uniform_flattening_stays_within_tolerance():
    sc = diff.random_scene(5, 128, 128, fills=8, cubics=4)
    st = sc.structure()
    v, k = sc.polygons(st)
    cub = diff._fill_cubics(sc.fills).detach().numpy()
    t = np.linspace(0, 1, 2001)
    B = np.stack([(1 - t) ** 3, 3 * t * (1 - t) ** 2, 3 * t ** 2 * (1 - t), t ** 3], 1)
    n = st.fill_steps
    base = 0
    for c in range(len(cub)):
        curve = B @ cub[c]
        # the chords of this cubic: its n points and the next cubic's first point
        idx = np.arange(n[c] + 1) / n[c]
        chord_pts = np.stack([np.interp(t, idx, (B[np.round(idx * 2000).astype(int)] @ cub[c])[:, j]) for j in (0, 1)], 1)
        assert np.max(np.linalg.norm(curve - chord_pts, axis=1)) <= diff.TOLERANCE + 1e-9
        base += n[c]


def test_recorded_lines_are_the_polygon_edges():
    sc = diff.random_scene(11, 120, 90, fills=6, strokes=4)
    v, k = sc.polygons()
    e = diff.encode(sc, v, k)  # raises when they differ
    assert e.scene.line_count == len(e.line_a) == int(e.per_draw.sum())
    assert e.scene.draw_count == len(k) == e.draw_total and e.scene.layer_count == 0
    assert e.per_draw.min() >= 3
    assert list(e.draw_ids) == list(range(len(k)))
    # with clip layers the clip draws take encoded draw indices of their own
    sc.clips = [(5.0, 5.0, 100.0, 80.0, 4.0, 2, 7)]
    e = diff.encode(sc, v, k)
    assert e.draw_total == len(k) + 1 and e.clip_ids == [2] and list(e.draw_ids[:3]) == [0, 1, 3]
    # a path clip's lines are recorded at its place in the stream and map to no vertex
    tri = np.array([[4.0, 4.0], [100.0, 10.0], [50.0, 76.0]])
    sc.clips = [(tri, 0, 1, 4)]
    e = diff.encode(sc, v, k)
    start = int(e.per_draw[0])
    assert e.scene.line_count == len(e.line_a) == int(e.per_draw.sum()) + 3
    assert list(e.line_a[start:start + 3]) == [-1, -1, -1] and (e.line_a[start + 3:] >= 0).all()
    assert e.draw_total == len(k) + 1 and e.clip_ids == [1]
    # a pixel-aligned rectangle becomes an SDF clip, with no lines of its own
    sc.clips = [(np.array([[8.0, 8.0], [90.0, 8.0], [90.0, 70.0], [8.0, 70.0]]), 0, 1, 4)]
    e = diff.encode(sc, v, k)
    assert e.scene.line_count == len(e.line_a) == int(e.per_draw.sum()) and (e.line_a >= 0).all()


def test_path_clip_masks_closed_forms():
    # An axis-aligned square clip from (4, 4) to (12.5, 10): whole pixels inside
    # take 1, the column cut at x = 12.5 takes one half, outside 0.  Traced twice
    # (winding 2) the inside keeps 1 under NonZero and becomes 0 under EvenOdd; the
    # cut column's area doubles to 1, which both rules take as 1.
    sq = np.array([[4.0, 4.0], [12.5, 4.0], [12.5, 10.0], [4.0, 10.0]])
    m = oracle.path_clip_coverage(sq, 0, 16, 16)
    assert np.allclose(m[4:10, 4:12], 1.0) and np.allclose(m[4:10, 12], 0.5)
    assert m.sum() == pytest.approx(8.5 * 6.0) and m[:4].sum() == 0 and m[10:].sum() == 0
    twice = np.concatenate([sq, sq])
    assert np.allclose(oracle.path_clip_coverage(twice, 0, 16, 16), np.where(m > 0, 1.0, 0.0))
    assert np.allclose(oracle.path_clip_coverage(twice, 1, 16, 16)[4:10, 4:12], 0.0)
    assert np.allclose(oracle.path_clip_coverage(twice, 1, 16, 16)[4:10, 12], 1.0)  # winding 1 over half


def test_stroke_outline_width_is_the_stroke_width():
    # A straight stroke: its outline is a capsule of the given width.
    strokes = torch.tensor([[[10.0, 20.0], [20.0, 20.0], [30.0, 20.0], [40.0, 20.0]]], dtype=torch.float64)
    sc = diff.DiffScene(64, 48, strokes=strokes, widths=torch.tensor([6.0], dtype=torch.float64),
                        colors=torch.tensor([[1.0, 0.0, 0.0, 1.0]], dtype=torch.float64))
    v, k = sc.polygons()
    p = v.detach().numpy()
    assert np.isclose(p[:, 1].max(), 23.0) and np.isclose(p[:, 1].min(), 17.0)
    assert np.isclose(p[:, 0].max(), 43.0) and np.isclose(p[:, 0].min(), 7.0)
    draws, lines, _, _ = diff.oracle_inputs(sc, v, k)
    image = oracle.diff_render(64, 48, [0, 0, 0, 0], draws, sc.premultiplied().detach().numpy(), lines)
    # capsule area: 30 x 6 rectangle + a disc of radius 3 (its polygon is slightly smaller)
    assert 30 * 6 + 0.97 * math.pi * 9 < image[:, :, 3].sum() <= 30 * 6 + math.pi * 9


def test_clipped_gradients_match_finite_differences():
    # Nested rounded-rect clips: the draws inside take the clips' coverage as a mask.
    sc = diff.random_scene(12, 96, 80, fills=6, strokes=3, cubics=3, radius_mu=math.log(14.0))
    sc.clips = [(10.5, 8.25, 80.0, 66.0, 7.0, 1, 8), (25.0, 12.5, 62.0, 50.0, 4.5, 3, 6)]
    w = weights(80, 96, 12)
    assert check_fd(sc, w) >= 30
    # the masks matter: the clipped image differs from the unclipped one
    with torch.no_grad():
        v, k = sc.polygons()
        draws, lines, _, _ = diff.oracle_inputs(sc, v, k)
        pm = sc.premultiplied().numpy()
        clipped = oracle.diff_render(96, 80, [0, 0, 0, 0], draws, pm, lines, masks=clip_masks(sc, v, k))
        plain = oracle.diff_render(96, 80, [0, 0, 0, 0], draws, pm, lines)
    assert np.abs(clipped - plain).max() > 0.1


def test_masked_oracle_equals_plain_oracle_under_unit_masks():
    sc = diff.random_scene(13, 64, 48, fills=5, strokes=2, radius_mu=math.log(12.0))
    with torch.no_grad():
        v, k = sc.polygons()
        draws, lines, _, _ = diff.oracle_inputs(sc, v, k)
        pm = sc.premultiplied().numpy()
    ones = [np.ones((48, 64))] * len(draws)
    g = weights(48, 64, 13)
    assert np.allclose(oracle.diff_render(64, 48, [0.1, 0, 0, 0.2], draws, pm, lines, masks=ones),
                       oracle.diff_render(64, 48, [0.1, 0, 0, 0.2], draws, pm, lines), atol=1e-12)
    la, ca = oracle.diff_backward(64, 48, [0.1, 0, 0, 0.2], draws, pm, lines, g, masks=ones)
    lb, cb = oracle.diff_backward(64, 48, [0.1, 0, 0, 0.2], draws, pm, lines, g)
    assert np.allclose(la, lb, atol=1e-10) and np.allclose(ca, cb, atol=1e-10)


def add_gradients(sc, draws, seed, kinds=None):
    """Gradient paints of 2-4 stops on `draws`, of kinds[draw] (linear when not
    given; a four-corner paint has 4)."""
    rng = np.random.default_rng(seed)
    kinds = kinds or {}
    for d in draws:
        kind = kinds.get(d, "linear")
        n = 4 if kind == "four_corner" else int(rng.integers(2, 5))
        x0, y0 = rng.uniform(0, sc.width), rng.uniform(0, sc.height)
        line = (x0, y0, x0 + rng.uniform(-60, 60), y0 + rng.uniform(-60, 60))
        offsets = np.sort(rng.uniform(0, 1, n))
        offsets[0] = min(offsets[0], 0.2)
        stops = torch.tensor(np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(0.4, 1.0, (n, 1))], 1),
                             dtype=sc.colors.dtype)
        sc.gradients[d] = (line, offsets, stops, kind)


def test_gradient_weights_reproduce_the_stops_and_interpolate():
    w = gradient_weights((10.0, 0.0, 50.0, 0.0), [0.0, 0.5, 1.0], 64, 4)
    assert np.allclose(w.sum(-1), 1.0)
    assert np.allclose(w[:, :10, 0], 1.0) and np.allclose(w[:, 50:, 2], 1.0)  # pad beyond the ends
    x = np.arange(64) + 0.5
    t = np.clip((x - 10) / 40, 0, 1)
    assert np.allclose(w[0, :, 1], np.where(t < 0.5, t / 0.5, (1 - t) / 0.5))  # the middle stop's tent


def test_radial_gradient_weights_on_integer_distances():
    # Centre (0.5, 0.5) at a pixel centre, radius 10: the pixel 3 right and 4 down
    # lies at distance 5 (t = 0.5) and takes the middle stop alone; 6 right and 8
    # down lies at 10 (t = 1) and takes the last; 2 along either axis (t = 0.2) is
    # 2/5 of the way from the first stop to the second; past the radius the end pads.
    w = gradient_weights((0.5, 0.5, 6.5, 8.5), [0.0, 0.5, 1.0], 16, 16, "radial")
    assert np.allclose(w.sum(-1), 1.0)
    assert np.allclose(w[4, 3], [0, 1, 0]) and np.allclose(w[8, 6], [0, 0, 1])
    assert np.allclose(w[0, 0], [1, 0, 0])
    assert np.allclose(w[0, 2], [1 - 2 / 5, 2 / 5, 0]) and np.allclose(w[2, 0], w[0, 2])
    assert np.allclose(w[15, 15], [0, 0, 1])
    # a radial paint is symmetric about its centre, unlike a linear one
    wl = gradient_weights((0.5, 0.5, 6.5, 8.5), [0.0, 0.5, 1.0], 16, 16)
    assert np.allclose(w[5, 2], w[2, 5]) and not np.allclose(wl[5, 2], wl[2, 5])


def test_gradient_kinds_in_paint_records():
    sc = diff.random_scene(3, 32, 32, fills=3)
    add_gradients(sc, [0, 2], 3, kinds={2: "radial"})
    rec = sc.paint_records().detach().numpy()
    assert rec[0, 0] == 1.0 and rec[1, 0] == 0.0 and rec[2, 0] == 2.0
    sc.gradients[1] = (*sc.gradients[0][:3], "conic")
    with pytest.raises(ValueError):
        sc.paint_records()
    sc.gradients[1] = (*sc.gradients[0][:3], "four_corner")  # needs 4 stops
    if len(sc.gradients[0][1]) != 4:
        with pytest.raises(ValueError):
            sc.paint_records()


def test_sweep_and_four_corner_weights_closed_forms():
    # Sweep about (8.5, 8.5) from +x: the pixel centre straight below (+y) lies a
    # quarter turn on (t = 0.25), the one to the left half a turn, straight above
    # three quarters; from a start pointing along +y everything turns back by 0.25.
    w = gradient_weights((8.5, 8.5, 20.0, 8.5), [0.0, 1.0], 17, 17, "sweep")
    assert np.allclose(w[12, 8], [0.75, 0.25]) and np.allclose(w[8, 2], [0.5, 0.5])
    assert np.allclose(w[2, 8], [0.25, 0.75]) and np.allclose(w[8, 12], [1.0, 0.0])
    w = gradient_weights((8.5, 8.5, 8.5, 30.0), [0.0, 1.0], 17, 17, "sweep")
    assert np.allclose(w[12, 8], [1.0, 0.0]) and np.allclose(w[8, 2], [0.75, 0.25])
    # Four corners over the box (0.5, 0.5)-(8.5, 4.5): each corner pixel takes its
    # stop alone; the box centre an equal quarter of each; beyond, the edges pad.
    w = gradient_weights((0.5, 0.5, 8.5, 4.5), None, 12, 8, "four_corner")
    assert np.allclose(w[0, 0], [1, 0, 0, 0]) and np.allclose(w[0, 8], [0, 1, 0, 0])
    assert np.allclose(w[4, 0], [0, 0, 1, 0]) and np.allclose(w[4, 8], [0, 0, 0, 1])
    assert np.allclose(w[2, 4], 0.25) and np.allclose(w[7, 11], [0, 0, 0, 1])
    assert np.allclose(w[1, 2], [0.75 * 0.75, 0.25 * 0.75, 0.75 * 0.25, 0.25 * 0.25])


@pytest.mark.parametrize("kind", ["radial", "sweep", "four_corner"])
def test_gradient_kind_stops_match_finite_differences(kind):
    sc = diff.random_scene(19, 96, 80, fills=6, strokes=1, cubics=3, radius_mu=math.log(16.0))
    add_gradients(sc, [0, 2, 4, 6], 19, kinds=dict.fromkeys([0, 2, 4, 6], kind))
    assert check_fd(sc, weights(80, 96, 19), per_param=10, rtol=1e-4) >= 30


def test_linear_gradient_stops_match_finite_differences():
    sc = diff.random_scene(16, 96, 80, fills=6, strokes=2, cubics=3, radius_mu=math.log(16.0))
    add_gradients(sc, [1, 3, 6], 16)
    # (stop gradients agree to ~1e-12; the bound is set by a stroke control point
    # near a join, where the geometry's curvature leaves 5e-5 in the differences)
    assert check_fd(sc, weights(80, 96, 16), per_param=10, rtol=1e-4) >= 30
