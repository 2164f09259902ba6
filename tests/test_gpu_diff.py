"""The differentiable form on the gfx950 device (tk_frame_render_diff,
tk_frame_backward) through the torch.autograd harness:

* the FP32 image against the FP64 oracle's exact-area composite;
* parameter gradients (control points, stroke widths, colours) against the FP64
  analytic gradient at the vertices rounded to FP32 (where the device renders:
  a badly conditioned entry moves by up to ~2e-3 relative between the FP64 and
  the FP32 positions) and against FP64 central finite differences (h = 1e-3 px),
  within 1e-3 relative;
* a batch of scenes against each scene alone;
* clipped scenes (SDF and path clip layers) against the same oracle with the
  clips' coverage as masks;
* scenes outside the form (SDF draws) are rejected."""
import math

import numpy as np
import pytest
import torch

import oracle
import tileink_amd as tk
from tileink_amd import diff

from test_diff import analytic_grads, clip_masks, fd_loss, paint_colors

pytestmark = pytest.mark.gpu

DEV = "cuda"


def gpu_grads(ctx, sc, w):
    params = [p for p in (sc.fills, sc.strokes, sc.widths, sc.colors) if p is not None]
    for p in params:
        p.grad = None
        p.requires_grad_(True)
    image = diff.render(ctx, sc, DEV)
    loss = (image.double() * torch.as_tensor(w, device=DEV)).sum()
    loss.backward()
    return image, [p.grad.clone() for p in params]


def oracle_image(sc):
    """The FP64 composite of the lines as recorded (FP32 vertices)."""
    with torch.no_grad():
        v, k = sc.polygons()
        draws, lines, _, _ = diff.oracle_inputs(sc, v, k, recorded=True)
        return oracle.diff_render(sc.width, sc.height, diff.clear_color(sc), draws,
                                  sc.premultiplied().detach().numpy(), lines, masks=clip_masks(sc, v, k))


def segment_image(sc):
    """The FP64 composite of the frame's segments (solid paints, no clips): each
    fill record's exact winding integral per pixel from the FP32 segments the
    oracle's path tiling writes (those the device writes, bit for bit), so that
    it differs from the device image by the device's own FP32 arithmetic only."""
    with torch.no_grad():
        v, k = sc.polygons()
    assert not sc.clips
    e = diff.encode(sc, v, k)
    _, out = oracle.render_frame(tk.SceneArrays(e.scene))
    ptcl, tw, th = out["ptcl"], e.scene.tiles_width, e.scene.tiles_height
    pix = np.arange(256, dtype=np.uint32)
    reqs, where = [], []
    for t in range(tw * th):
        for r in range(out["tile_ptcl_starts"][t], out["tile_ptcl_ends"][t]):
            rec = ptcl[r]
            if rec[0] != 1:  # TK_PTCL_FILL
                continue
            q = np.zeros((256, 8), np.uint32)
            q[:, 0], q[:, 1], q[:, 2] = rec[3], rec[4], rec[1]
            q[:, 3], q[:, 4] = pix % 16, pix // 16
            reqs.append(q)
            where.append((t, int(rec[5]), int(rec[2])))
    cov = oracle.fill_coverage_f64(np.concatenate(reqs), out["segments"]).reshape(-1, 16, 16)
    colors = sc.premultiplied().detach().numpy()
    image = np.empty((th * 16, tw * 16, 4))
    image[:] = diff.clear_color(sc)
    for (t, d, rule), c in zip(where, cov):
        a = np.abs(c - 2.0 * np.rint(0.5 * c)) if rule == 1 else np.minimum(np.abs(c), 1.0)
        block = image[(t // tw) * 16:(t // tw) * 16 + 16, (t % tw) * 16:(t % tw) * 16 + 16]
        block[:] = block * (1.0 - a[..., None] * colors[d][3]) + a[..., None] * colors[d]
    return image[:sc.height, :sc.width]


def assert_grads_close(got, ref, rtol=1e-3, floor=0.02):
    """|got - ref| <= rtol * max(|ref|, floor * max|ref|) per parameter."""
    for g, r in zip(got, ref):
        g, r = g.cpu().double().numpy().ravel(), r.cpu().double().numpy().ravel()
        scale = np.abs(r).max()
        bound = rtol * np.maximum(np.abs(r), floor * scale)
        bad = np.abs(g - r) > bound + 1e-12
        assert not bad.any(), (int(bad.sum()), g[bad][:5], r[bad][:5])


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_forward_matches_fp64_oracle(ctx, seed):
    sc = diff.random_scene(seed, 160, 112, fills=24, strokes=8, radius_mu=math.log(18.0))
    image = diff.render(ctx, sc, DEV).cpu().double().numpy()
    ref = oracle_image(sc)
    assert image.shape == ref.shape
    err = float(np.abs(image - ref).max())
    assert err < 1e-5, err


@pytest.mark.parametrize("seed", [3, 6])
def test_gradients_match_fp64_analytic(ctx, seed):
    sc = diff.random_scene(seed, 160, 112, fills=24, strokes=8, radius_mu=math.log(18.0))
    w = np.random.default_rng(seed).normal(size=(112, 160, 4))
    _, got = gpu_grads(ctx, sc, w)
    analytic_grads(sc, w, recorded=True)
    ref = [p.grad for p in (sc.fills, sc.strokes, sc.widths, sc.colors)]
    assert_grads_close(got, ref)


def test_gradients_with_opaque_overlaps_and_even_odd(ctx):
    sc = diff.random_scene(9, 96, 96, fills=10, strokes=4, cubics=3, radius_mu=math.log(20.0))
    with torch.no_grad():
        sc.colors[::2, 3] = 1.0
    sc.rules[:] = 1
    sc.clear = (0.2, 0.1, 0.0, 0.4)
    w = np.random.default_rng(9).normal(size=(96, 96, 4))
    image, got = gpu_grads(ctx, sc, w)
    err = float(np.abs(image.detach().cpu().double().numpy() - oracle_image(sc)).max())
    assert err < 1e-5, err
    analytic_grads(sc, w, recorded=True)
    assert_grads_close(got, [p.grad for p in (sc.fills, sc.strokes, sc.widths, sc.colors)])


@pytest.mark.parametrize("seed", [24, 25])
def test_path_clipped_scenes_match_fp64(ctx, seed):
    # Path clip layers (closed polygons under either fill rule: the draws' alpha is
    # multiplied by the clip's FP32 exact-area alpha), nested with an SDF clip.
    sc = diff.random_scene(seed, 160, 112, fills=20, strokes=6, radius_mu=math.log(18.0))
    star = np.array([[80 + 50 * math.cos(a) * (1.0 if i % 2 == 0 else 0.45),
                      56 + 50 * math.sin(a) * (1.0 if i % 2 == 0 else 0.45)]
                     for i, a in enumerate(np.linspace(0, 2 * math.pi, 11)[:-1] + 0.3)])
    # a self-overlapping pentagram: its centre has winding 2 (kept by NonZero, a hole under EvenOdd)
    penta = np.array([[40 + 30 * math.cos(a), 60 + 30 * math.sin(a)]
                      for a in np.linspace(0, 4 * math.pi, 6)[:-1] + 0.1])
    sc.clips = [(star, 0, 2, 20), (12.5, 9.25, 140.0, 100.0, 9.0, 6, 14), (penta, 1, 21, 26)]
    w = np.random.default_rng(seed).normal(size=(112, 160, 4))
    image, got = gpu_grads(ctx, sc, w)
    err = float(np.abs(image.detach().cpu().double().numpy() - oracle_image(sc)).max())
    assert err < 1e-5, err
    analytic_grads(sc, w, recorded=True)
    assert_grads_close(got, [p.grad for p in (sc.fills, sc.strokes, sc.widths, sc.colors)])


@pytest.mark.parametrize("seed", [14, 15])
def test_clipped_scenes_match_fp64(ctx, seed):
    # Nested rounded-rect clip layers around ranges of draws (SDF clips: the draws'
    # alpha is multiplied by the clips' FP32 coverage).
    sc = diff.random_scene(seed, 160, 112, fills=20, strokes=6, radius_mu=math.log(18.0))
    sc.clips = [(12.5, 9.25, 140.0, 100.0, 9.0, 2, 20), (40.0, 20.5, 110.0, 80.0, 5.5, 6, 14),
                (0.0, 0.0, 60.0, 112.0, 0.0, 22, 26)]
    w = np.random.default_rng(seed).normal(size=(112, 160, 4))
    image, got = gpu_grads(ctx, sc, w)
    err = float(np.abs(image.detach().cpu().double().numpy() - oracle_image(sc)).max())
    assert err < 1e-5, err
    analytic_grads(sc, w, recorded=True)
    assert_grads_close(got, [p.grad for p in (sc.fills, sc.strokes, sc.widths, sc.colors)])


@pytest.mark.parametrize("seed", [17, 18])
def test_gradient_paints_match_fp64(ctx, seed):
    # Linear, radial, sweep and four-corner gradient paints (2-4 stops) on some
    # draws, inside a clip: the image and the gradients of geometry, colours and
    # stop colours against the FP64 oracle.
    from test_diff import add_gradients
    sc = diff.random_scene(seed, 160, 112, fills=16, strokes=4, radius_mu=math.log(18.0))
    add_gradients(sc, [0, 3, 5, 8, 11, 13, 16, 18], seed,
                  kinds={3: "radial", 8: "radial", 5: "sweep", 16: "sweep", 11: "four_corner", 18: "four_corner"})
    sc.clips = [(10.0, 8.0, 150.0, 104.0, 8.0, 2, 18)]
    w = np.random.default_rng(seed).normal(size=(112, 160, 4))
    stops = [g[2] for g in sc.gradients.values()]
    for p in stops:
        p.requires_grad_(True)
    image, got = gpu_grads(ctx, sc, w)
    got_stops = [p.grad.clone() for p in stops]
    with torch.no_grad():
        v, k = sc.polygons()
        draws, lines, _, _ = diff.oracle_inputs(sc, v, k, recorded=True)
        ref_image = oracle.diff_render(160, 112, diff.clear_color(sc), draws, paint_colors(sc), lines,
                                       masks=clip_masks(sc, v, k))
    err = float(np.abs(image.detach().cpu().double().numpy() - ref_image).max())
    assert err < 1e-5, err
    analytic_grads(sc, w, recorded=True)
    assert_grads_close(got + got_stops, [p.grad for p in (sc.fills, sc.strokes, sc.widths, sc.colors)] +
                       [p.grad for p in stops])


def test_gradients_match_finite_differences_at_1024(ctx):
    # 1024^2, 2048 paths of 4 cubics (1536 filled, 512 stroked), loss against a
    # target whose control points are perturbed by N(0, 4 px) and colours by
    # N(0, 0.05); central differences in FP64 (h = 1e-3 px) on 64 random
    # parameters: fill and stroke control points, stroke widths and colours.
    sc = diff.random_scene(41, 1024, 1024, fills=1536, strokes=512, cubics=4, stroke_cubics=4)
    tgt = diff.random_scene(41, 1024, 1024, fills=1536, strokes=512, cubics=4, stroke_cubics=4)
    rng = np.random.default_rng(42)
    with torch.no_grad():
        tgt.fills += torch.as_tensor(rng.normal(0, 4, tuple(tgt.fills.shape)))
        tgt.strokes += torch.as_tensor(rng.normal(0, 4, tuple(tgt.strokes.shape)))
        tgt.colors += torch.as_tensor(rng.normal(0, 0.05, tuple(tgt.colors.shape)))
        tgt.colors.clamp_(0, 1)
        target = diff.render(ctx, tgt, DEV).double()
    image = diff.render(ctx, sc, DEV).detach().cpu().double().numpy()
    # The device's own arithmetic: within 1e-5 of the FP64 composite of the same
    # segments.
    err = float(np.abs(image - segment_image(sc)).max())
    assert err < 1e-5, err
    # The segments themselves: path tiling clips each line at tile edges in
    # absolute FP32 coordinates, where one ulp near x = 1024 is 6.1e-5 px, so the
    # segments' coverage differs from the lines' exact areas by up to ~4.5e-5 per
    # draw, and a pixel under several translucent layers sums those (this scene:
    # 2.94e-4 between the FP64 composites of its segments and of its lines).
    err = float(np.abs(image - oracle_image(sc)).max())
    assert err < 5e-4, err
    named = (("fills", sc.fills), ("strokes", sc.strokes), ("widths", sc.widths), ("colors", sc.colors))
    for _, p in named:
        p.requires_grad_(True)
    image = diff.render(ctx, sc, DEV)
    loss = ((image.double() - target) ** 2).sum()
    loss.backward()
    w = 2.0 * (image.detach().double() - target).cpu().numpy()  # dL/dimage, fixed for the differences
    st = sc.structure()
    h = 1e-3
    picks = [("fills", sc.fills, i) for i in rng.choice(sc.fills.numel(), 24, replace=False)] + \
            [("strokes", sc.strokes, i) for i in rng.choice(sc.strokes.numel(), 16, replace=False)] + \
            [("widths", sc.widths, i) for i in rng.choice(sc.widths.numel(), 8, replace=False)] + \
            [("colors", sc.colors, i) for i in rng.choice(sc.colors.numel(), 16, replace=False)]
    scale = {n: float(p.grad.abs().max()) for n, p in named}
    for name, p, i in picks:
        flat = p.data.view(-1)
        o = float(flat[i])
        flat[i] = o + h
        lp = fd_loss(sc, st, w)
        flat[i] = o - h
        lm = fd_loss(sc, st, w)
        flat[i] = o
        fd = (lp - lm) / (2 * h)
        g = float(p.grad.view(-1)[i])
        assert abs(g - fd) <= 1e-3 * max(abs(fd), 0.02 * scale[name]), (name, i, g, fd)


def test_batch_equals_each_scene_alone(ctx):
    scenes = [diff.random_scene(20 + i, w, h, fills=f, strokes=s, radius_mu=math.log(12.0))
              for i, (w, h, f, s) in enumerate([(64, 48, 12, 3), (200, 90, 30, 6), (33, 130, 8, 2)])]
    scenes[1].clips = [(20.5, 10.0, 180.0, 80.0, 6.0, 5, 30)]
    scenes[2].clips = [(np.array([[2.0, 5.0], [31.0, 20.0], [16.5, 120.0]]), 0, 1, 9)]  # a path clip
    ws = [np.random.default_rng(i).normal(size=(s.height, s.width, 4)) for i, s in enumerate(scenes)]
    singles = []
    for sc, w in zip(scenes, ws):
        image, grads = gpu_grads(ctx, sc, w)
        singles.append((image.detach().cpu().numpy(), grads))
    for sc in scenes:
        for p in (sc.fills, sc.strokes, sc.widths, sc.colors):
            p.grad = None
    images = diff.render(ctx, scenes, DEV)
    loss = sum((im.double() * torch.as_tensor(w, device=DEV)).sum() for im, w in zip(images, ws))
    loss.backward()
    for sc, im, (ref_im, ref_g) in zip(scenes, images, singles):
        assert np.array_equal(im.detach().cpu().numpy(), ref_im)
        # Line gradients are summed in a fixed order; colour gradients with float
        # atomics, so they agree to rounding.
        assert_grads_close([p.grad for p in (sc.fills, sc.strokes, sc.widths, sc.colors)], ref_g, rtol=1e-4)


def test_scenes_with_sdf_draws_are_rejected(ctx):
    # SDF draws (a rounded rectangle painted directly, SDF records) are not in the form.
    canvas = tk.Canvas(64, 64)
    canvas.push_polygon(np.array([[10.0, 10.0], [50.0, 12.0], [30.0, 50.0]]), tk.Color(1, 0, 0, 1))
    canvas.push_rect(8.0, 8.0, 40.0, 30.0, 4.0, tk.Color(0, 1, 0, 1))
    scene = canvas.encode(0)
    frame = tk.Frame(ctx, scene)
    colors = torch.ones((scene.draw_count, 32), device=DEV)
    image = torch.zeros((64 * 64, 4), device=DEV)
    with pytest.raises(tk.TileinkError):
        tk._abi.check(tk._abi.lib().tk_frame_render_diff(frame._h, colors.data_ptr(), image.data_ptr()))
    grad = torch.zeros_like(image)
    with pytest.raises(tk.TileinkError):  # no differentiable render ran
        tk._abi.check(tk._abi.lib().tk_frame_backward(frame._h, colors.data_ptr(), grad.data_ptr(),
                                                      image.data_ptr(), colors.data_ptr()))


def test_backward_needs_its_own_forward(ctx):
    # The tile pass replaces the forward's kept coverages with their gradients, and a
    # plain render replaces the PTCL: either way another backward needs another
    # differentiable render and is refused until then, never answered from stale state.
    sc = diff.random_scene(11, 64, 48, fills=4, cubics=3, radius_mu=math.log(10.0))
    with torch.no_grad():
        v, k = sc.polygons()
        plan = diff.Plan(ctx, [sc], [v], [k])
        colors = plan.full_paint(sc.paint_records().to(DEV)).to(torch.float32).contiguous()
    frame = plan.frame
    image = torch.zeros((plan.pixels, 4), device=DEV)
    grad = torch.ones_like(image)
    lines = torch.zeros((max(plan.line_count, 1), 4), device=DEV)
    cgrad = torch.zeros_like(colors)
    lib = tk._abi.lib()
    args = (frame._h, colors.data_ptr(), grad.data_ptr(), lines.data_ptr(), cgrad.data_ptr())
    # The native calls run on the context's stream: wait for torch's fills before
    # them and for the context before reading what they wrote.
    torch.cuda.synchronize()
    tk._abi.check(lib.tk_frame_render_diff(frame._h, colors.data_ptr(), image.data_ptr()))
    tk._abi.check(lib.tk_frame_backward(*args))
    ctx.synchronize()
    first = (lines.clone(), cgrad.clone())
    with pytest.raises(tk.TileinkError):
        tk._abi.check(lib.tk_frame_backward(*args))
    tk._abi.check(lib.tk_frame_render_diff(frame._h, colors.data_ptr(), image.data_ptr()))
    frame.render()
    with pytest.raises(tk.TileinkError):
        tk._abi.check(lib.tk_frame_backward(*args))
    torch.cuda.synchronize()
    tk._abi.check(lib.tk_frame_render_diff(frame._h, colors.data_ptr(), image.data_ptr()))
    tk._abi.check(lib.tk_frame_backward(*args))
    ctx.synchronize()
    assert torch.equal(lines, first[0])
    assert torch.allclose(cgrad, first[1], rtol=1e-5, atol=1e-6)


def test_adam_fitting_recovers_a_perturbed_scene(ctx):
    # Image fitting: control points and colours perturbed from a target scene are
    # optimised with Adam against the target's image; the loss falls.
    target = diff.random_scene(31, 96, 80, fills=8, cubics=4, dtype=torch.float32, radius_mu=math.log(14.0))
    target.clips = [(4.0, 4.0, 92.0, 76.0, 6.0, 0, 8)]
    sc = diff.random_scene(31, 96, 80, fills=8, cubics=4, dtype=torch.float32, radius_mu=math.log(14.0))
    sc.clips = list(target.clips)
    rng = np.random.default_rng(32)
    with torch.no_grad():
        ref = diff.render(ctx, target, DEV).clone()
        sc.fills += torch.as_tensor(rng.normal(0, 2.0, tuple(sc.fills.shape)), dtype=torch.float32)
        sc.colors += torch.as_tensor(rng.normal(0, 0.05, tuple(sc.colors.shape)), dtype=torch.float32)
        sc.colors.clamp_(0, 1)
    sc.fills.requires_grad_(True)
    sc.colors.requires_grad_(True)
    opt = torch.optim.Adam([sc.fills, sc.colors], lr=0.05)
    losses = []
    for _ in range(40):
        opt.zero_grad()
        loss = ((diff.render(ctx, sc, DEV) - ref) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            sc.colors.clamp_(0, 1)
        losses.append(float(loss.detach()))
    assert losses[-1] < 0.5 * losses[0], losses


def _c5_like(count, seed, size=256, lo=64, hi=513):
    """C5's shape at `count` scenes: size x size, lo..hi-1 closed paths of 4 cubics,
    10 % gradient paints (half radial), one rounded-rectangle clip each."""
    rng = np.random.default_rng(seed)
    paths = rng.integers(lo, hi, count)
    scenes = []
    for i, n in enumerate(paths):
        sc = diff.random_scene(seed * 1000 + i, size, size, fills=int(n), cubics=4, radius_mu=math.log(16.0))
        inset = rng.uniform(4, 24, 4)
        sc.clips = [(inset[0], inset[1], size - inset[2], size - inset[3], float(rng.uniform(4, 16)), 0, int(n))]
        diff.random_gradients(sc, 0.1, seed * 7000 + i, radial=0.5)
        scenes.append(sc)
    return scenes


def test_renderer_moves_geometry_through_the_gpu_front_end(ctx):
    # The second and later renders of a diff.Renderer send only the new vertices to
    # tk_frame_set_polygons; images and gradients equal a fresh host recording of
    # the same parameters (a path clip's lines map to no vertex).
    scenes = _c5_like(5, 3, size=160, lo=20, hi=60)
    scenes[2].clips = [(np.array([[2.0, 5.0], [150.0, 20.0], [80.5, 150.0]]), 1, 0, len(scenes[2].order))]
    r = diff.Renderer(ctx, scenes, DEV)
    r()
    rng = np.random.default_rng(4)
    for step in range(3):
        with torch.no_grad():
            for sc in scenes:
                sc.fills += torch.as_tensor(rng.normal(0, 1.5, tuple(sc.fills.shape)))
        ws = [torch.as_tensor(np.random.default_rng(step * 10 + i).normal(size=(s.height, s.width, 4)), device=DEV)
              for i, s in enumerate(scenes)]
        results = []
        for make in (lambda: r(), lambda: diff.render(ctx, scenes, DEV)):
            params = [p for s in scenes for p in (s.fills, s.colors, *[g[2] for g in s.gradients.values()])]
            for p in params:
                p.grad = None
                p.requires_grad_(True)
            images = make()
            sum((im.double() * w).sum() for im, w in zip(images, ws)).backward()
            results.append(([im.detach().cpu() for im in images], [p.grad.clone() for p in params]))
            for p in params:
                p.requires_grad_(False)
        (ia, ga), (ib, gb) = results
        assert all(torch.equal(a, b) for a, b in zip(ia, ib)), step
        # colour and stop gradients are summed with float atomics: equal to rounding
        assert_grads_close(ga, gb, rtol=1e-4)


def test_c5_batch_gradients_match_fp64_on_sampled_scenes(ctx):
    # 256 scenes of 256x256 with 64-512 paths each, gradient paints and one clip:
    # the batched forward + backward, and on three sampled scenes every parameter
    # gradient against the FP64 analytic gradient.
    scenes = _c5_like(256, 5)
    ws = [np.random.default_rng(100 + i).normal(size=(256, 256, 4)) for i in range(len(scenes))]
    params = [[p for p in (s.fills, s.colors, *[g[2] for g in s.gradients.values()])] for s in scenes]
    for ps in params:
        for p in ps:
            p.grad = None
            p.requires_grad_(True)
    images = diff.Renderer(ctx, scenes, DEV)()
    sum((im.double() * torch.as_tensor(w, device=DEV)).sum() for im, w in zip(images, ws)).backward()
    got = {i: [p.grad.clone() for p in params[i]] for i in range(len(scenes))}
    sampled = [0, 97, 255]
    assert any(scenes[i].gradients for i in sampled)
    for i in sampled:
        sc = scenes[i]
        for p in params[i]:
            p.grad = None
        analytic_grads(sc, ws[i], recorded=True)
        assert_grads_close(got[i], [p.grad for p in params[i]])


def test_recomputed_and_kept_coverage_give_the_same_gradients(ctx):
    # The backward recomputes every (record, pixel) coverage from the segments by
    # default; with diff_keep_coverage the forward stores them instead.  The two
    # give the same image and the same line gradients bit for bit (the coverage
    # sums are exact, so recomputing reproduces them), colour gradients to the
    # rounding of their float atomics.
    scenes = _c5_like(6, 9, size=192, lo=30, hi=90)
    scenes[1].clips = [(np.array([[3.0, 4.0], [180.0, 30.0], [90.5, 185.0]]), 0, 0, len(scenes[1].order))]
    polys = [s.polygons() for s in scenes]
    plan = diff.Plan(ctx, scenes, [v for v, _ in polys], [k for _, k in polys])
    colors = plan.full_paint(torch.cat([s.paint_records() for s in scenes])).to(DEV, torch.float32)
    grad = torch.as_tensor(np.random.default_rng(3).normal(size=(plan.pixels, 4)), dtype=torch.float32, device=DEV)
    out = {}
    for keep in (0, 1):
        plan.frame.set_tuning(diff_keep_coverage=keep)
        image = plan.forward(colors).clone()
        lines, cols = plan.backward(grad)
        out[keep] = (image, lines.clone(), cols.clone())
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
    # Float atomics add a draw's per-wave sums in arrival order, so an entry whose
    # terms cancel carries the rounding of partial sums as large as the largest
    # entries: compare on that scale.
    diff_max = float((out[0][2] - out[1][2]).abs().max())
    scale = float(out[1][2].abs().max())
    assert diff_max <= 1e-5 * scale, (diff_max, scale)


def test_stream_ordered_plans_match_waiting_plans(ctx):
    # Plans with sync=False order their calls against torch's stream only; several
    # of them issued back to back give the images and gradients of plans that wait
    # for the device after every call.
    scenes = _c5_like(4, 11, size=128, lo=16, hi=48)
    polys = [s.polygons() for s in scenes]
    results = {}
    for sync in (True, False):
        plans = [diff.Plan(ctx, [s], [v], [k], sync=sync) for s, (v, k) in zip(scenes, polys)]
        outs = []
        for p, s in zip(plans, scenes):
            colors = p.full_paint(s.paint_records()).to(DEV, torch.float32)
            image = p.forward(colors)
            lines, cols = p.backward(2.0 * image)
            outs.append((image, lines, cols))
        for p in plans:
            p.check_status()
        results[sync] = [(a.cpu(), b.cpu(), c.cpu()) for a, b, c in outs]
    for (ia, la, ca), (ib, lb, cb) in zip(results[True], results[False]):
        assert torch.equal(ia, ib)
        assert torch.equal(la, lb)
        assert float((ca - cb).abs().max()) <= 1e-5 * float(cb.abs().max())
