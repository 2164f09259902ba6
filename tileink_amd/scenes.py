"""Seeded synthetic scenes for the benchmark and parity workloads.

Each generator records a scene through the Canvas API, the way an application
would, from a fixed seed (0x7117E1C0 + workload number).  Geometry, paint and
layer schedules are fixed by the seed, so the oracle and the gfx950 pipeline
see identical inputs.
"""
import math

import numpy as np

from . import (PATTERN_BILINEAR, PATTERN_NEAREST, BezPath, Canvas, Color, FillRule, Image, ImageBrush,
               LinearGradient)

SEED_BASE = 0x7117E1C0


def _rng(workload, salt=0):
    return np.random.Generator(np.random.PCG64(SEED_BASE + workload + 1000003 * salt))


def _color(rng, alpha_lo=0.3):
    r, g, b = rng.uniform(0.0, 1.0, 3)
    return Color(r, g, b, rng.uniform(alpha_lo, 1.0))


def _cubic_path(rng, n_cubics, lo, hi, closed):
    """A contour of n cubic segments with points uniform in the box [lo, hi]."""
    pts = rng.uniform(lo, hi, size=(3 * n_cubics + 1, 2))
    p = BezPath().move_to(*pts[0])
    for i in range(n_cubics):
        a, b, c = pts[3 * i + 1], pts[3 * i + 2], pts[3 * i + 3]
        if closed and i == n_cubics - 1:
            c = pts[0]
        p.curve_to(a[0], a[1], b[0], b[1], c[0], c[1])
    if closed:
        p.close_path()
    return p


def _gradient(rng, lo, hi):
    n = int(rng.integers(2, 5))
    offsets = np.sort(rng.uniform(0.0, 1.0, n))
    offsets[0], offsets[-1] = 0.0, 1.0
    start = rng.uniform(lo, hi)
    end = rng.uniform(lo, hi)
    return LinearGradient(tuple(start), tuple(end), [(float(o), _color(rng)) for o in offsets],
                          extend=int(rng.integers(0, 3)))


def mixed_scene_640x360(width=640, height=360):
    """640x360: 256 rounded rects, 256 filled closed cubic paths, 64 stroked paths, shuffled draw order."""
    rng = _rng(1)
    canvas = Canvas(width, height, 1.0)
    kinds = np.array([0] * 256 + [1] * 256 + [2] * 64)
    rng.shuffle(kinds)
    size = np.array([width, height], dtype=np.float64)
    for kind in kinds:
        if kind == 0:
            c = rng.uniform(0.0, 1.0, 2) * size
            w, h = rng.uniform(8.0, 160.0, 2)
            canvas.push_rect(c[0] - w / 2, c[1] - h / 2, c[0] + w / 2, c[1] + h / 2, float(rng.uniform(0.0, 24.0)),
                             _color(rng))
        elif kind == 1:
            path = _cubic_path(rng, int(rng.integers(4, 13)), (0.0, 0.0), tuple(size), True)
            rule = FillRule.EvenOdd if rng.uniform() < 0.5 else FillRule.NonZero
            canvas.push_path(path, _color(rng), rule=rule, tolerance=0.25)
        else:
            path = _cubic_path(rng, int(rng.integers(4, 13)), (0.0, 0.0), tuple(size), True)
            width_px = float(rng.lognormal(math.log(2.0), 0.5))
            canvas.push_stroke(path, width_px, _color(rng), tolerance=0.25)
    return canvas


def synthetic_image(rng, height, width, opaque=False):
    """A straight-alpha RGBA8 image of smooth ramps plus noise, premultiplied."""
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    phase = rng.uniform(0.0, 6.3, 4)
    base = [127.5 + 127.5 * np.sin(x / max(width, 1) * (3 + k) + y / max(height, 1) * (2 + k) + phase[k])
            for k in range(4)]
    rgba = np.stack(base, -1) + rng.normal(0.0, 12.0, size=(height, width, 4))
    rgba = np.clip(np.rint(rgba), 0, 255).astype(np.uint8)
    if opaque:
        rgba[..., 3] = 255
    return Image.from_rgba8(rgba)


def image_scene(width=320, height=240, scale=1.0, seed=0):
    """Images drawn every way the canvas offers: image rects with fractional
    edges (bilinear pad and nearest), image brushes on rounded rects, filled and
    stroked paths under rotations (repeat and reflect, bilinear and nearest,
    opacity below 255), a clip layer around an image, over a gradient background."""
    rng = _rng(20, seed)
    canvas = Canvas(width, height, scale)
    size = np.array([width, height], dtype=np.float64)
    canvas.push_rect(0.0, 0.0, width, height, 0.0, _gradient(rng, (0.0, 0.0), tuple(size)))
    photo = canvas.add_image(synthetic_image(rng, 47, 61, opaque=True))
    icon = canvas.add_image(synthetic_image(rng, 9, 7))
    strip = canvas.add_image(synthetic_image(rng, 3, 40))
    canvas.push_image(0.11 * width + 0.25, 0.05 * height + 0.5, 0.62 * width + 0.75, 0.55 * height + 0.1, photo)
    canvas.push_image(0.7 * width, 0.04 * height, 0.95 * width, 0.3 * height, icon, sampling=PATTERN_NEAREST)
    canvas.push_image(0.05 * width, 0.8 * height, 0.9 * width, 0.86 * height, strip, extend=1)
    for k in range(12):
        c = rng.uniform(0.0, 1.0, 2) * size
        angle = rng.uniform(0.0, 2 * math.pi)
        cos, sin = math.cos(angle), math.sin(angle)
        cell = rng.uniform(6.0, 30.0)
        # user space -> a rotated and scaled cell of the unit square
        xf = (cos / cell, -sin / cell, sin / cell, cos / cell, rng.uniform(-1, 1), rng.uniform(-1, 1))
        brush = ImageBrush([photo, icon, strip][k % 3], xf, extend=int(rng.integers(0, 3)),
                           sampling=[PATTERN_BILINEAR, PATTERN_NEAREST][k % 2],
                           opacity=255 if k % 4 else int(rng.integers(60, 250)))
        if k % 3 == 0:
            w, h = rng.uniform(20.0, 0.5 * width, 2)
            canvas.push_rect(c[0] - w / 2, c[1] - h / 2, c[0] + w / 2, c[1] + h / 2, float(rng.uniform(0.0, 16.0)),
                             brush)
        elif k % 3 == 1:
            path = _cubic_path(rng, int(rng.integers(3, 8)), (0.0, 0.0), tuple(size), True)
            canvas.push_path(path, brush, transform=(cos, sin, -sin, cos, 0.0, 0.0) if k % 2 else None,
                             rule=FillRule.EvenOdd if k % 2 else FillRule.NonZero)
        else:
            path = _cubic_path(rng, int(rng.integers(3, 8)), (0.0, 0.0), tuple(size), False)
            canvas.push_stroke(path, float(rng.uniform(3.0, 14.0)), brush)
    clip = _cubic_path(rng, 6, (0.2 * width, 0.2 * height), (0.8 * width, 0.8 * height), True)
    canvas.push_clip_layer(clip)
    canvas.push_image(0.0, 0.0, width, height, photo, extend=2)
    canvas.push_rect(0.3 * width, 0.3 * height, 0.5 * width, 0.6 * height, 6.0, _color(rng))
    canvas.pop_layer()
    return canvas


def clipped_paths_scene(width=4096, height=4096, paths=200_000, clips=32, workload=2):
    """4096^2, 200k paths (60% fill, 40% stroke), 4-16 cubics each, 10% linear gradients,
    `clips` nested rounded-rect clip layers pushed then popped between equal draw groups."""
    rng = _rng(workload)
    canvas = Canvas(width, height, 1.0)
    groups = 2 * clips + 1
    per_group = np.full(groups, paths // groups)
    per_group[: paths - per_group.sum()] += 1
    clip_box = np.array([0.0, 0.0, float(width), float(height)])
    drawn = 0
    for g in range(groups):
        for _ in range(int(per_group[g])):
            anchor = rng.uniform(0.0, 1.0, 2) * (width, height)
            extent = float(rng.lognormal(math.log(48.0), 1.0))
            lo = anchor - extent / 2
            hi = anchor + extent / 2
            n = int(rng.integers(4, 17))
            fill = rng.uniform() < 0.6
            brush = _gradient(rng, lo, hi) if rng.uniform() < 0.1 else _color(rng)
            if fill:
                rule = FillRule.EvenOdd if rng.uniform() < 0.5 else FillRule.NonZero
                canvas.push_path(_cubic_path(rng, n, lo, hi, True), brush, rule=rule, tolerance=0.25)
            else:
                w = float(rng.lognormal(math.log(1.5), 0.7))
                canvas.push_stroke(_cubic_path(rng, n, lo, hi, False), w, brush, tolerance=0.25)
            drawn += 1
        if g < clips:
            # Each clip sits inside the previous one, each side pulled in by up to 3% of the current size.
            w, h = clip_box[2] - clip_box[0], clip_box[3] - clip_box[1]
            m = rng.uniform(0.0, 0.03, 4) * (w, h, w, h)
            clip_box = clip_box + (m[0], m[1], -m[2], -m[3])
            canvas.push_clip_sdf_rect_layer(*clip_box, float(rng.uniform(0.0, 64.0)))
        elif g < 2 * clips:
            canvas.pop_layer()
    return canvas


def clipped_paths_small(scale=8, paths=None):
    """clipped-paths generator on a (4096/scale)^2 canvas with proportionally fewer paths (parity cases)."""
    side = 4096 // scale
    n = paths if paths is not None else max(64, 200_000 // (scale * scale))
    return clipped_paths_scene(side, side, n, clips=min(32, max(2, n // 64)), workload=20 + scale)


def scene_set_sizes(count, seed=3, path_mu=math.log(1500.0), path_sigma=1.2, max_paths=20_000):
    """(width, height, paths) of `count` independent scenes: side log-uniform in
    [64, 2048] px, aspect ratio uniform in [0.5, 2], paths lognormal(path_mu,
    path_sigma) clipped to [10, max_paths]."""
    rng = _rng(seed)
    out = []
    for _ in range(count):
        side = math.exp(rng.uniform(math.log(64.0), math.log(2048.0)))
        aspect = rng.uniform(0.5, 2.0)
        w = int(min(max(round(side * math.sqrt(aspect)), 16), 4096))
        h = int(min(max(round(side / math.sqrt(aspect)), 16), 4096))
        n = int(min(max(round(rng.lognormal(path_mu, path_sigma)), 10), max_paths))
        out.append((w, h, n))
    return out


def scene_set_member(width, height, paths, index):
    """One scene of the independent-scene workload: the clipped-paths generator at this
    size, with clip nesting scaled to the path count."""
    return clipped_paths_scene(width, height, paths, clips=min(32, max(1, paths // 256)), workload=1000 + index)


def _encoded_member(job):
    from . import SceneArrays
    index, (w, h, n) = job
    return index, SceneArrays(scene_set_member(w, h, n, index).encode(0))


def scene_set(sizes, workers=None):
    """Records and encodes every scene of `sizes` (scene_set_sizes) in worker
    processes; returns their SceneArrays in order.  Each scene comes from its own
    seed, so the result does not depend on the worker count.  The workers load
    the recorder only (they start fresh, not forked from a process that may hold
    the GPU)."""
    import multiprocessing as mp
    import os
    if workers is None:
        workers = min(12, len(os.sched_getaffinity(0)))
    jobs = list(enumerate(sizes))
    if workers <= 1:
        return [_encoded_member(j)[1] for j in jobs]
    out = [None] * len(jobs)
    # Largest scenes first, so that the pool ends evenly.
    jobs.sort(key=lambda j: -j[1][2])
    with mp.get_context("spawn").Pool(workers) as pool:
        for index, arrays in pool.imap_unordered(_encoded_member, jobs, chunksize=1):
            out[index] = arrays
    return out
