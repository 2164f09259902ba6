"""The differentiable form: gradients of the FP32 coverage and composite with
respect to control points, stroke widths and paint colours.

The GPU side is two C-ABI entry points (tk_frame_render_diff, tk_frame_backward):
the forward composites the scene's FILL records in FP32 without the RGBA8 round
trips of the plain renderer, and the backward returns dL/d(line endpoints) and
dL/d(colour) for a given dL/d(image).  The geometry that turns parameters into
lines is written here in torch, so autograd carries the line gradients back to
the parameters:

  * a fill is a closed contour of cubics; each cubic is cut at uniform parameter
    steps t = k/n, with n fixed from the control polygon so that the chord error
    stays under `tolerance` (|B''| <= 6 max|P0 - 2P1 + P2|, error <= |B''|/(8 n^2));
  * a stroke is an open chain of cubics, flattened the same way; its outline is
    one closed polygon: the left offsets, a round end cap, the right offsets
    backwards and a round start cap.  At a vertex turning by at most 120 degrees
    each side gets the miter point q + h m, m = (n0 + n1) / (1 + n0.n1); a
    sharper turn gets both segments' offset points (a bevel).

The structure of the geometry (subdivision counts, miter or bevel, cap points)
is fixed from the parameters' values and held constant under differentiation,
as it is for any flattening.  Scenes may hold fills and strokes only (no
layers), and the plain renderer's u8 quantization is not part of this form.
"""
import ctypes as C
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

import tileink_amd as tk
from tileink_amd import _abi
from tileink_amd._abi import check, lib

TOLERANCE = 0.25
_MAX_SUBDIVISIONS = 256


def _bernstein(t):
    mt = 1.0 - t
    return torch.stack([mt * mt * mt, 3.0 * t * mt * mt, 3.0 * t * t * mt, t * t * t], -1)


def _subdivisions(cubics, tolerance):
    """Uniform steps per cubic ((N, 4, 2) array) for a chord error under tolerance.
    The second differences' norms are sqrt(x * x + y * y) in FP64, term by term as
    _subdivisions_on_device evaluates them, so both give the same counts."""
    p = np.asarray(cubics, np.float64)
    a = p[:, 0] - 2 * p[:, 1] + p[:, 2]
    b = p[:, 1] - 2 * p[:, 2] + p[:, 3]
    dd = np.maximum(np.sqrt(a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]), np.sqrt(b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1]))
    return np.clip(np.ceil(np.sqrt(0.75 * dd / tolerance)), 1, _MAX_SUBDIVISIONS).astype(np.int64)


def _subdivisions_on_device(cubics, tolerance):
    """_subdivisions of a (N, 4, 2) tensor, evaluated where it lives (FP64, the same
    operations in the same order); only the counts come back to the host."""
    p = cubics.detach().double()
    a = p[:, 0] - 2 * p[:, 1] + p[:, 2]
    b = p[:, 1] - 2 * p[:, 2] + p[:, 3]
    dd = torch.maximum(torch.sqrt(a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]), torch.sqrt(b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1]))
    if np.ndim(tolerance):
        tolerance = torch.as_tensor(tolerance, dtype=torch.float64, device=p.device)
    steps = torch.clamp(torch.ceil(torch.sqrt(0.75 * dd / tolerance)), 1, _MAX_SUBDIVISIONS)
    return steps.to(torch.int64).cpu().numpy()


def _flatten_index(steps, closed_ends, dtype, device):
    """(cubic of each point, its t) for _flatten, as device tensors."""
    m = steps + closed_ends.astype(np.int64)  # points per cubic
    ci = np.repeat(np.arange(len(steps)), m)
    k = np.arange(int(m.sum())) - np.repeat(np.cumsum(m) - m, m)
    t = torch.as_tensor(k / np.repeat(steps, m), dtype=dtype, device=device)
    return torch.as_tensor(ci, device=device), t


def _flatten(cubics, steps, closed_ends, index=None):
    """Points of cubics (M, 4, 2) at t = k/n, k < n, plus t = 1 where closed_ends
    marks a chain's last cubic (open chains end on their last point).  index: the
    _flatten_index of these steps, when the caller keeps it."""
    ci, t = index if index is not None else _flatten_index(steps, closed_ends, cubics.dtype, cubics.device)
    w = _bernstein(t)
    sel = cubics[ci]
    return (w[:, :, None] * sel).sum(1)


class Structure:
    """The discrete choices of one geometry evaluation (kept fixed for its gradient)."""

    def __init__(self):
        self.fill_steps = None
        self.stroke_steps = None
        self.stroke_keep = None
        self.recipe = None


class DiffScene:
    """Parameters of a differentiable scene.

    fills:   (F, C, 3, 2) — per closed contour, per cubic: first control point,
             second control point, end point (a cubic starts at the previous end).
    strokes: (S, 3K + 1, 2) — open chains of K cubics; widths (S,).
    colors:  (F + S, 4) straight-alpha RGBA per draw, in painter order.
    order:   painter order as (kind, index) pairs, kind 0 fill / 1 stroke
             (default: all fills, then all strokes).
    rules:   fill rule per fill (0 NonZero, 1 EvenOdd); strokes use NonZero.
    clips:   clip layers around the draws [first, end) of painter order (ranges
             nest or are disjoint): SDF layers as (x0, y0, x1, y1, radius, first,
             end), a rounded rectangle; path layers as (points (n, 2), rule,
             first, end), a closed polygon under fill rule `rule`.  The clip
             geometry is not differentiated.
    gradients: {painter-order draw: (line (x0, y0, x1, y1), offsets (n,),
             stops (n, 4) straight RGBA tensor[, kind])} -- gradient paints of
             1-4 stops, kind "linear" (the default: t = the projection on the
             line), "radial" (t = the distance from the line's start over its
             length), "sweep" (t = the angle about the line's start from its
             direction, over a full turn) or "four_corner" (4 stops at the
             corners of the box the line spans, bilinear; offsets unused); the
             stop colours are parameters, the line and offsets are not.  Such a
             draw's entry in `colors` is unused.
    """

    def __init__(self, width, height, fills=None, strokes=None, widths=None, colors=None, order=None, rules=None,
                 clear=(0.0, 0.0, 0.0, 0.0), tolerance=TOLERANCE, clips=None, gradients=None):
        self.width, self.height = int(width), int(height)
        self.fills = fills
        self.strokes = strokes
        self.widths = widths
        self.colors = colors
        nf = 0 if fills is None else int(fills.shape[0])
        ns = 0 if strokes is None else int(strokes.shape[0])
        self.order = list(order) if order is not None else [(0, i) for i in range(nf)] + [(1, i) for i in range(ns)]
        self.rules = np.zeros(nf, np.int64) if rules is None else np.asarray(rules, np.int64)
        self.clear = tuple(float(c) for c in clear)
        self.tolerance = float(tolerance)
        self.clips = [tuple(c) for c in (clips or [])]
        self.gradients = dict(gradients or {})

    # --- geometry -------------------------------------------------------------------

    def structure(self):
        st = Structure()
        if self.fills is not None and self.fills.shape[0]:
            st.fill_steps = _subdivisions(_fill_cubics(self.fills).detach().cpu().numpy(), self.tolerance)
        if self.strokes is not None and self.strokes.shape[0]:
            self._stroke_structure(st)
        return st

    def _stroke_structure(self, st):
        cub = _stroke_cubics(self.strokes)
        st.stroke_steps = _subdivisions(cub.detach().cpu().numpy(), self.tolerance)
        q = _flatten(cub.detach().double(), st.stroke_steps, _chain_ends(self.strokes)).cpu().numpy()
        st.stroke_keep, st.recipe, st.stroke_kept = _stroke_recipe(
            q, _stroke_counts(self.strokes, st.stroke_steps), self.widths.detach().cpu().numpy() * 0.5,
            self.tolerance)

    def polygons(self, st=None):
        """(vertices (V, 2), vertices per draw in painter order) as torch tensors of
        the parameters' dtype; differentiable in the parameters."""
        st = st or self.structure()
        dtype = (self.fills if self.fills is not None else self.strokes).dtype
        device = (self.fills if self.fills is not None else self.strokes).device
        parts, part_counts = [], [np.zeros(0, np.int64), np.zeros(0, np.int64)]
        if st.fill_steps is not None:
            cub = _fill_cubics(self.fills)
            parts.append(_flatten(cub, st.fill_steps, np.zeros(len(st.fill_steps), bool)))
            part_counts[0] = st.fill_steps.reshape(self.fills.shape[0], -1).sum(1)
        if st.stroke_steps is not None:
            cub = _stroke_cubics(self.strokes)
            q = _flatten(cub, st.stroke_steps, _chain_ends(self.strokes))
            outlines = _stroke_outlines(q, st.stroke_keep, st.recipe, st.stroke_kept, 0.5 * self.widths)
            parts.append(torch.cat(outlines) if outlines else torch.zeros((0, 2), dtype=dtype, device=device))
            part_counts[1] = np.array([len(r) for r in st.recipe], np.int64)
        if not parts:
            return torch.zeros((0, 2), dtype=dtype, device=device), np.zeros(0, np.int64)
        # painter order: one gather over fills then strokes laid end to end
        allv = torch.cat(parts) if len(parts) > 1 else parts[0]
        fill_vertices = int(part_counts[0].sum())
        first_of = [np.cumsum(part_counts[0]) - part_counts[0],
                    fill_vertices + np.cumsum(part_counts[1]) - part_counts[1]]
        gather, counts = self._painter_gather(part_counts, first_of)
        verts = allv.index_select(0, torch.as_tensor(gather, device=allv.device))
        return verts, counts

    def _painter_gather(self, part_counts, first_of):
        """(indices of the painter-order vertices among the fill and stroke vertices,
        vertices per draw) from each fill's and stroke's vertex count and first
        vertex; (kind, index) arrays of the order are kept while it is unchanged."""
        key = (id(self.order), len(self.order))
        if getattr(self, "_order_key", None) != key:
            self._order_kinds = np.array([kd for kd, _ in self.order], np.int64)
            self._order_index = np.array([i for _, i in self.order], np.int64)
            self._order_key = key
        kinds, idx = self._order_kinds, self._order_index
        counts = np.zeros(len(idx), np.int64)
        first = np.zeros(len(idx), np.int64)
        for kind in (0, 1):
            sel = kinds == kind
            if sel.any():
                counts[sel] = part_counts[kind][idx[sel]]
                first[sel] = first_of[kind][idx[sel]]
        gather = np.arange(int(counts.sum())) - np.repeat(np.cumsum(counts) - counts, counts) + np.repeat(first, counts)
        return gather, counts

    def draw_rules(self):
        return np.array([int(self.rules[i]) if kind == 0 else 0 for kind, i in self.order], np.uint32)

    def premultiplied(self, colors=None):
        c = self.colors if colors is None else colors
        return torch.cat([c[:, :3] * c[:, 3:4], c[:, 3:4]], 1)

    def paint_records(self):
        """(draws, 32) paint records in painter order (the layout tk_frame_render_diff
        reads), differentiable in the colours and the gradient stops."""
        return batch_paint_records([self])

    def _paint_layout(self):
        """The fixed part of the paint records: (head (draws, 16): kind, line,
        stop count and offsets per draw; each gradient's draw; per stop, its
        gradient and slot; the stop tensors in gradient order).  Kept while the
        gradient entries and the draw count are unchanged."""
        n = int(self.colors.shape[0])
        items = sorted(self.gradients.items())
        key = (n, tuple((d, id(entry)) for d, entry in items))
        if getattr(self, "_layout_key", None) != key:
            head = np.zeros((n, 16), np.float64)
            head[:, 5] = 1.0
            rows, which, slot = [], [], []
            for g, (d, entry) in enumerate(items):
                line, offsets, st, kind = gradient_parts(entry)
                k = len(offsets)
                if not 1 <= k <= 4 or st.shape[0] != k or (kind == "four_corner" and k != 4):
                    raise ValueError("a gradient has 1-4 stops (a four-corner one 4), one colour each")
                head[d, 0] = GRADIENT_KINDS[kind]
                head[d, 1:5] = line
                head[d, 5] = k
                head[d, 6:6 + k] = offsets
                rows.append(d)
                which.extend([g] * k)
                slot.extend(range(k))
            self._layout = (head, np.asarray(rows, np.int64), np.asarray(which, np.int64),
                            np.asarray(slot, np.int64))
            self._layout_key = key
            self._layout_entries = [entry for _, entry in items]  # held, so their ids stay theirs
        return self._layout + ([gradient_parts(entry)[2] for _, entry in items],)


GRADIENT_KINDS = {"linear": 1.0, "radial": 2.0, "sweep": 3.0, "four_corner": 4.0}


def gradient_parts(entry):
    """(line, offsets, stops, kind) of a gradient entry, kind "linear" when not given."""
    if len(entry) == 3:
        return (*entry, "linear")
    line, offsets, stops, kind = entry
    if kind not in GRADIENT_KINDS:
        raise ValueError(f"unknown gradient kind {kind!r}")
    return line, offsets, stops, kind


# --- the geometry and paint of many scenes at once ------------------------------------
#
# A batch of scenes (an optimiser step over hundreds of them) evaluates each step
# of the geometry once for all scenes: the fills of every scene flatten in one
# pass (one host copy of the subdivision counts, one set of device operations),
# and every scene's paint records come from one premultiplication and one scatter
# of all gradient stops.  Elementwise, the values are those of the per-scene
# methods bit for bit.


def _joined(tensors):
    """torch.cat(tensors) along dim 0; when they are contiguous slices laid end to
    end over all of one tensor (a caller keeping many scenes' parameters in one
    tensor), that tensor itself, so autograd reaches it without a copy and a
    split per scene."""
    if len(tensors) == 1:
        return tensors[0]
    base = tensors[0]._base
    if base is not None and base.is_contiguous() and all(t._base is base and t.is_contiguous() for t in tensors):
        pos = base.storage_offset()
        for t in tensors:
            if t.storage_offset() != pos:
                break
            pos += t.numel()
        else:
            if pos == base.storage_offset() + base.numel():
                return base.reshape(-1, *tensors[0].shape[1:])
    return torch.cat(tensors)


def _fill_groups(scenes):
    """Indices of the scenes with fills, grouped by cubics per contour."""
    groups = {}
    for i, s in enumerate(scenes):
        if s.fills is not None and s.fills.shape[0]:
            groups.setdefault(int(s.fills.shape[1]), []).append(i)
    return groups


def batch_structures(scenes):
    """DiffScene.structure of every scene, the fill subdivisions of all scenes from
    one host copy per cubic count."""
    structs = [Structure() for _ in scenes]
    for c, idx in _fill_groups(scenes).items():
        fills = _joined([scenes[i].fills for i in idx]).detach()
        sizes = [int(scenes[i].fills.shape[0]) * c for i in idx]
        tols = np.array([scenes[i].tolerance for i in idx], np.float64)
        tol = float(tols[0]) if (tols == tols[0]).all() else np.repeat(tols, sizes)
        steps = _subdivisions_on_device(_fill_cubics(fills), tol)
        ends = np.cumsum(sizes)
        for i, lo, hi in zip(idx, ends - sizes, ends):
            structs[i].fill_steps = steps[lo:hi]
    for s, st in zip(scenes, structs):
        if s.strokes is not None and s.strokes.shape[0]:
            s._stroke_structure(st)
    return structs


def batch_polygons(scenes, structs=None, cache=None):
    """DiffScene.polygons of every scene as one tensor: (vertices (V, 2) of all
    scenes laid end to end in painter order, vertices per draw of each scene, the
    first vertex of each scene and the total); differentiable in the parameters.
    cache: a dict the caller keeps between calls on the same scenes (diff.Renderer
    does); the index tensors are built again only when the subdivision counts
    change."""
    structs = structs or batch_structures(scenes)
    cache = {} if cache is None else cache
    parts, base = [], 0
    none = np.zeros(0, np.int64)
    part_counts = [[none, none] for _ in scenes]
    first_of = [[none, none] for _ in scenes]
    signature = [np.array([[id(s.order), len(s.order)] for s in scenes], np.int64)]
    for c, idx in _fill_groups(scenes).items():
        cub = _fill_cubics(_joined([scenes[i].fills for i in idx]))
        steps = np.concatenate([structs[i].fill_steps for i in idx])
        signature.append(steps)
        hit = cache.get(("flatten", c))
        if hit is None or not np.array_equal(hit[0], steps) or hit[1].device != cub.device or hit[2].dtype != cub.dtype:
            hit = (steps, *_flatten_index(steps, np.zeros(len(steps), bool), cub.dtype, cub.device))
            cache[("flatten", c)] = hit
        parts.append(_flatten(cub, steps, None, index=hit[1:]))
        for i in idx:
            per = structs[i].fill_steps.reshape(scenes[i].fills.shape[0], -1).sum(1)
            part_counts[i][0] = per
            first_of[i][0] = base + np.cumsum(per) - per
            base += int(per.sum())
    for i, (s, st) in enumerate(zip(scenes, structs)):
        if st.stroke_steps is None:
            continue
        q = _flatten(_stroke_cubics(s.strokes), st.stroke_steps, _chain_ends(s.strokes))
        outlines = _stroke_outlines(q, st.stroke_keep, st.recipe, st.stroke_kept, 0.5 * s.widths)
        per = np.array([len(r) for r in st.recipe], np.int64)
        if outlines:
            parts.append(torch.cat(outlines))
        part_counts[i][1] = per
        first_of[i][1] = base + np.cumsum(per) - per
        base += int(per.sum())
        signature.append(per)
    ref = next((t for s in scenes for t in (s.fills, s.strokes) if t is not None), None)
    hit = cache.get("gather")
    if (hit is None or len(hit[0]) != len(signature) or not all(np.array_equal(a, b) for a, b in zip(hit[0], signature))
            or hit[1].device != ref.device):
        gathers, counts, firsts = [], [], [0]
        for i, s in enumerate(scenes):
            g, k = s._painter_gather(part_counts[i], first_of[i])
            gathers.append(g)
            counts.append(k)
            firsts.append(firsts[-1] + int(k.sum()))
        gather = torch.as_tensor(np.concatenate(gathers) if gathers else none, device=ref.device)
        hit = ([a.copy() for a in signature], gather, counts, np.asarray(firsts, np.int64))
        cache["gather"] = hit
    _, gather, counts, firsts = hit
    if not parts:
        return torch.zeros((0, 2), dtype=ref.dtype, device=ref.device), counts, firsts
    allv = torch.cat(parts) if len(parts) > 1 else parts[0]
    return allv.index_select(0, gather), counts, firsts


def batch_paint_records(scenes):
    """DiffScene.paint_records of every scene laid end to end: one premultiplication
    of all colours and one scatter of all gradient stops; differentiable."""
    colors = _joined([s.colors for s in scenes])
    pm = torch.cat([colors[:, :3] * colors[:, 3:4], colors[:, 3:4]], 1)
    n = pm.shape[0]
    heads, rows, which, slot, stop_list = [], [], [], [], []
    d0 = g0 = 0
    for s in scenes:
        head, r, w, sl, st = s._paint_layout()
        heads.append(head)
        rows.append(r + d0)
        which.append(w + g0)
        slot.append(sl)
        stop_list.extend(st)
        d0 += head.shape[0]
        g0 += len(r)
    head = np.concatenate(heads) if len(heads) > 1 else heads[0]
    stops = torch.cat([pm, pm.new_zeros((n, 12))], 1)
    if stop_list:
        st = _joined(stop_list).to(device=pm.device, dtype=pm.dtype)
        spm = torch.cat([st[:, :3] * st[:, 3:4], st[:, 3:4]], 1)
        dev = pm.device
        blocks = pm.new_zeros((g0, 4, 4)).index_put(
            (torch.as_tensor(np.concatenate(which), device=dev), torch.as_tensor(np.concatenate(slot), device=dev)),
            spm)
        stops = stops.index_put((torch.as_tensor(np.concatenate(rows), device=dev),), blocks.reshape(-1, 16))
    return torch.cat([torch.as_tensor(head, dtype=pm.dtype, device=pm.device), stops], 1)


def _fill_cubics(fills):
    ends = fills[:, :, 2]
    starts = torch.roll(ends, 1, dims=1)
    return torch.stack([starts, fills[:, :, 0], fills[:, :, 1], ends], 2).reshape(-1, 4, 2)


def _stroke_cubics(strokes):
    S, n = strokes.shape[0], (strokes.shape[1] - 1) // 3
    idx = torch.arange(n, device=strokes.device)[:, None] * 3 + torch.arange(4, device=strokes.device)
    return strokes[:, idx].reshape(S * n, 4, 2)


def _chain_ends(strokes):
    n = (strokes.shape[1] - 1) // 3
    ends = np.zeros((strokes.shape[0], n), bool)
    ends[:, -1] = True
    return ends.reshape(-1)


def _stroke_counts(strokes, steps):
    n = (strokes.shape[1] - 1) // 3
    return steps.reshape(strokes.shape[0], n).sum(1) + 1


def _cap_steps(h, tol):
    step = 2.0 * math.acos(1.0 - tol / h) if h > tol else math.pi
    return max(1, int(math.ceil(math.pi / step)))


def _stroke_recipe(q, counts, half_widths, tol):
    """Per stroke outline point: (centreline vertex, segment, a, b, c) for
    q[v] + h (a n[seg] + b u[seg] + c m[v]); vertex and segment indices are
    into the kept centreline points and the stroke's segments, laid end to end."""
    keep = np.zeros(len(q), bool)
    recipe, kept_counts = [], []
    start = 0
    seg_base = 0
    kept_base = 0
    for s, cnt in enumerate(counts):
        pts = q[start:start + cnt]
        k = [0]
        for i in range(1, cnt):
            if np.hypot(*(pts[i] - pts[k[-1]])) > 1e-9:
                k.append(i)
        keep[start + np.array(k)] = True
        m = len(k) - 1  # segments
        if m < 1:
            raise ValueError(f"stroke {s} has no segment of nonzero length")
        kept_counts.append(len(k))
        rows = []
        if m >= 1:
            p = pts[k]
            d = p[1:] - p[:-1]
            u = d / np.linalg.norm(d, axis=1, keepdims=True)
            n = np.stack([-u[:, 1], u[:, 0]], 1)
            V, G = kept_base, seg_base
            miter = [False] * (m + 1)
            for i in range(1, m):
                miter[i] = 1.0 + float(n[i - 1] @ n[i]) >= 0.5
            left = [(V + 0, G + 0, 1.0, 0.0, 0.0)]
            for i in range(1, m):
                if miter[i]:
                    left.append((V + i, G + i, 0.0, 0.0, 1.0))
                else:
                    left += [(V + i, G + i - 1, 1.0, 0.0, 0.0), (V + i, G + i, 1.0, 0.0, 0.0)]
            left.append((V + m, G + m - 1, 1.0, 0.0, 0.0))
            kc = _cap_steps(half_widths[s], 0.25 * tol)
            end_cap = [(V + m, G + m - 1, math.cos(math.pi * j / kc), math.sin(math.pi * j / kc), 0.0)
                       for j in range(1, kc)]
            right = [(V + m, G + m - 1, -1.0, 0.0, 0.0)]
            for i in range(m - 1, 0, -1):
                if miter[i]:
                    right.append((V + i, G + i, 0.0, 0.0, -1.0))
                else:
                    right += [(V + i, G + i, -1.0, 0.0, 0.0), (V + i, G + i - 1, -1.0, 0.0, 0.0)]
            right.append((V + 0, G + 0, -1.0, 0.0, 0.0))
            start_cap = [(V + 0, G + 0, -math.cos(math.pi * j / kc), -math.sin(math.pi * j / kc), 0.0)
                         for j in range(1, kc)]
            rows = left + end_cap + right + start_cap
        recipe.append(np.array(rows, np.float64).reshape(-1, 5))
        start += cnt
        seg_base += m
        kept_base += len(k)
    return keep, recipe, np.array(kept_counts, np.int64)


def _stroke_outlines(q, keep, recipe, kept_counts, half_widths):
    """Outline polygons of the strokes from the recipe; differentiable in q and h."""
    counts = [len(r) for r in recipe]
    kept = q[torch.as_tensor(np.nonzero(keep)[0], device=q.device)]
    # segments: consecutive kept points of one stroke
    seg_from, seg_to = [], []
    offs = 0
    for nk in kept_counts:
        seg_from += list(range(offs, offs + nk - 1))
        seg_to += list(range(offs + 1, offs + nk))
        offs += int(nk)
    dev = q.device
    a_i = torch.as_tensor(seg_from, device=dev)
    b_i = torch.as_tensor(seg_to, device=dev)
    d = kept[b_i] - kept[a_i]
    u = d / torch.linalg.norm(d, dim=1, keepdim=True)
    n = torch.stack([-u[:, 1], u[:, 0]], 1)
    # miter vectors at every kept vertex (zero where unused): m = (n_prev + n_next) / (1 + n_prev . n_next)
    m = torch.zeros_like(kept)
    prev_seg = np.full(len(kept), -1)
    next_seg = np.full(len(kept), -1)
    prev_seg[np.asarray(seg_to)] = np.arange(len(seg_to))
    next_seg[np.asarray(seg_from)] = np.arange(len(seg_from))
    inner = np.nonzero((prev_seg >= 0) & (next_seg >= 0))[0]
    if len(inner):
        ip = torch.as_tensor(prev_seg[inner], device=dev)
        inx = torch.as_tensor(next_seg[inner], device=dev)
        num = n[ip] + n[inx]
        den = 1.0 + (n[ip] * n[inx]).sum(1, keepdim=True)
        m = m.index_put((torch.as_tensor(inner, device=dev),), num / den)
    rec = np.concatenate([r for r in recipe if len(r)])
    stroke_of_point = np.repeat(np.arange(len(recipe)), counts)
    vi = torch.as_tensor(rec[:, 0].astype(np.int64), device=dev)
    si = torch.as_tensor(rec[:, 1].astype(np.int64), device=dev)
    coef = torch.as_tensor(rec[:, 2:5], dtype=q.dtype, device=dev)
    h = half_widths[torch.as_tensor(stroke_of_point, device=dev)][:, None]
    pts = kept[vi] + h * (coef[:, 0:1] * n[si] + coef[:, 1:2] * u[si] + coef[:, 2:3] * m[vi])
    return list(torch.split(pts, counts))


# --- scene encoding and the line <-> vertex map ----------------------------------------


def edge_map(verts32, counts):
    """The lines the recorder emits for closed polygons: edge j of a polygon runs
    from vertex j to vertex j + 1 (the last back to the first); edges shorter than
    1e-12 are dropped, as push_line drops them.  Returns (line -> first vertex,
    line -> second vertex, lines per draw)."""
    counts = np.asarray(counts, np.int64)
    ia = np.arange(int(counts.sum()), dtype=np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]) if len(counts) else np.zeros(0, np.int64)
    ib = ia + 1
    ends = starts + counts - 1
    ib[ends[counts > 0]] = starts[counts > 0]
    p = np.asarray(verts32, np.float32).astype(np.float64)
    d = p[ib] - p[ia] if len(ia) else np.zeros((0, 2))
    ok = (d * d).sum(1) > 1.0e-24
    poly = np.repeat(np.arange(len(counts)), counts)
    ok &= np.repeat(counts, counts) > 1
    per_draw = np.bincount(poly[ok], minlength=len(counts)).astype(np.int64)
    return ia[ok], ib[ok], per_draw


class Encoded:
    """A scene recorded from polygons: the canvas and encoded scene, the line ->
    vertex map (line_a, line_b), lines per painter-order draw, each painter-order
    draw's index among the encoded draws (draw_ids), the encoded draw count and
    the encoded index of each clip layer's draw (clip_ids)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def encode(scene, verts, counts, colors=None):
    """Records the polygons (inside the scene's clip layers) on a canvas and
    encodes it; checks that the recorded lines are the polygon edges and the
    path clips' edges."""
    v32 = verts.detach().to("cpu", torch.float32).numpy()
    canvas = tk.Canvas(scene.width, scene.height)
    rules = scene.draw_rules()
    col = (scene.colors if colors is None else colors).detach().cpu().double().numpy()
    clips = scene.clips
    # outer clips first where several open at one draw, inner first where several close
    opens = sorted(range(len(clips)), key=lambda i: (clips[i][-2], -clips[i][-1]))
    counts = np.asarray(counts, np.int64)
    offsets = np.concatenate([[0], np.cumsum(counts)])
    colors32 = np.clip(col, 0.0, 1.0).astype(np.float32)
    la, lb, per_draw = edge_map(v32, counts)
    line_of = np.concatenate([[0], np.cumsum(per_draw)])
    # The expected lines in recording order, as (p0|p1 rows, line -> vertex maps);
    # a path clip's lines map to no vertex (-1).
    pieces = []
    # what each recorded path is, in path order: ("clip", clip index) or ("draws", lo, hi)
    sources = []
    # runs of draws between clip boundaries go down in one call each
    bounds = sorted({0, len(counts)} | {int(c[-2]) for c in clips} | {int(c[-1]) for c in clips})
    draw_ids, o, open_stack = [], 0, []
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        while o < len(opens) and clips[opens[o]][-2] == lo:
            c = clips[opens[o]]
            if len(c) == 7:
                canvas.push_clip_sdf_rect_layer(*c[:5])
            else:
                pts = np.asarray(c[0], np.float64).reshape(-1, 2)
                path = tk.BezPath().move_to(*pts[0])
                for q in pts[1:]:
                    path.line_to(*q)
                canvas.push_clip_layer(path.close_path(), rule=tk.FillRule(int(c[1])))
                p32 = pts.astype(np.float32)
                ca, cb, _ = edge_map(p32, [len(p32)])
                none = np.full(len(ca), -1, np.int64)
                pieces.append((np.concatenate([p32[ca], p32[cb]], 1), none, none, opens[o]))
                sources.append(("clip", opens[o]))
            open_stack.append(opens[o])
            o += 1
        if hi > lo:
            draw_ids.extend(canvas.push_polygons(v32[offsets[lo]:offsets[hi]], counts[lo:hi], colors32[lo:hi],
                                                 rules[lo:hi]))
            a, b = la[line_of[lo]:line_of[hi]], lb[line_of[lo]:line_of[hi]]
            pieces.append((np.concatenate([v32[a], v32[b]], 1), a, b, None))
            sources.append(("draws", lo, hi))
        while open_stack and clips[open_stack[-1]][-1] == hi:
            canvas.pop_layer()
            open_stack.pop()
    if o != len(opens) or open_stack:
        raise ValueError("clip ranges must nest, lie within the draws and not be empty")
    clear = np.clip(np.asarray(scene.clear), 0, 1)
    cw = int(round(clear[0] * 255)) | int(round(clear[1] * 255)) << 8 | int(round(clear[2] * 255)) << 16 | \
        int(round(clear[3] * 255)) << 24
    enc = canvas.encode(cw)
    records = _abi.view(enc.draws, _abi.DrawRecord, enc.draw_count)
    clip_draws = [int(i) for i in np.nonzero(records["tag"] == _abi.DRAW_CLIP)[0]]
    clip_ids = [0] * len(clips)
    for n, i in enumerate(opens):
        clip_ids[i] = clip_draws[n]
    # a path clip the recorder took as an SDF clip (a pixel-aligned rectangle) has no lines
    pieces = [p for p in pieces if p[3] is None or records["sdf_offset"][clip_ids[p[3]]] == _abi.INVALID]
    sources = [c for c in sources if c[0] != "clip" or records["sdf_offset"][clip_ids[c[1]]] == _abi.INVALID]
    lines = _abi.view(enc.lines, _abi.LINE_DTYPE, enc.line_count)
    got = np.concatenate([lines["p0"], lines["p1"]], 1) if len(lines) else np.zeros((0, 4), np.float32)
    want = np.concatenate([p[0] for p in pieces]) if pieces else np.zeros((0, 4), np.float32)
    if got.shape != want.shape or not np.array_equal(got, want):
        raise tk.TileinkError(_abi.TK_ERR_INVALID, "recorded lines do not match the polygon edges")
    if any(len(c) != 7 for c in clips):
        la = np.concatenate([p[1] for p in pieces]) if pieces else la
        lb = np.concatenate([p[2] for p in pieces]) if pieces else lb
    return Encoded(canvas=canvas, scene=enc, line_a=la, line_b=lb, per_draw=per_draw,
                   draw_ids=np.asarray(draw_ids, np.int64), draw_total=int(enc.draw_count), clip_ids=clip_ids,
                   sources=sources)


def clear_color(scene):
    """The clear colour the frame starts from: each channel of scene.clear taken
    as premultiplied and rounded to 8 bits, as encode packs it."""
    return np.round(np.clip(np.asarray(scene.clear, np.float64), 0, 1) * 255.0) / 255.0


# --- the autograd function ------------------------------------------------------------


class Plan:
    """One encoded scene (or batch of scenes) on the device, for forward and backward.

    sync=True: each forward and backward waits for the device and forward reads the
    render's status flags (raising on an overflow).  sync=False: the calls are only
    ordered against torch's current stream (stream waits, no host wait), and the
    flags are read by check_status(), for loops that issue many plans back to back."""

    def __init__(self, context, scenes, verts, counts, colors=None, sync=True):
        self.context = context
        self.scenes = scenes
        self.sync = sync
        self._stream = None
        self.encoded, self.canvases, la, lb, self.sizes = [], [], [], [], []
        v_off, l_off = 0, 0
        self.line_ranges = []
        colors_of = colors or [None] * len(scenes)
        draw_index, d_off = [], 0
        self.encodings = []
        jobs = list(zip(scenes, verts, counts, colors_of))
        if len(jobs) > 1:
            # Scenes record on their own canvases; the native calls release the
            # interpreter lock, so the recordings run side by side.
            with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
                encodings = list(pool.map(lambda job: encode(*job), jobs))
        else:
            encodings = [encode(*job) for job in jobs]
        for (s, v, k, col), e in zip(jobs, encodings):
            self.encodings.append(e)
            self.canvases.append(e.canvas)
            self.encoded.append(e.scene)
            la.append(np.where(e.line_a < 0, -1, e.line_a + v_off))
            lb.append(np.where(e.line_b < 0, -1, e.line_b + v_off))
            self.line_ranges.append(e.per_draw)
            draw_index.append(e.draw_ids + d_off)
            v_off += len(v)
            d_off += e.draw_total
            self.sizes.append((s.width, s.height, e.scene.tiles_width * e.scene.tiles_height))
        self.draw_index = np.concatenate(draw_index)
        self.vertex_count = v_off
        if len(scenes) == 1:
            self.frame = tk.Frame(context, self.encoded[0])
            self._batch = None
            self.offsets = [0]
            self.pixels = scenes[0].width * scenes[0].height
        else:
            self._batch = tk.Batch(context, self.encoded)
            self.frame = self._batch.frame
            cols = np.cumsum([0] + [t for _, _, t in self.sizes])
            self.offsets = [int(c) * 256 for c in cols[:-1]]
            self.pixels = int(cols[-1]) * 256
        # path clip lines (-1) take their gradient into a sink row past the vertices
        self.line_a = np.concatenate(la)
        self.line_b = np.concatenate(lb)
        self.line_a[self.line_a < 0] = v_off
        self.line_b[self.line_b < 0] = v_off
        self._la_t = self._lb_t = None
        self.line_count = len(self.line_a)
        self.draw_count = d_off  # encoded draws, clip draws included

    def _index(self, device):
        if self._la_t is None or self._la_t.device != device:
            src_a = self.line_a if self.line_a is not None else self._la_t
            src_b = self.line_b if self.line_b is not None else self._lb_t
            self._la_t = torch.as_tensor(src_a, device=device)
            self._lb_t = torch.as_tensor(src_b, device=device)
        return self._la_t, self._lb_t

    def _clip_points(self, dev):
        """Every path clip's points (FP32, in the scenes' source order) on the device,
        and the first point of each (scene, clip)."""
        if getattr(self, "_clip_pts", None) is None or self._clip_pts.device != dev:
            pts, first = [], {}
            n = 0
            for i, (sc, e) in enumerate(zip(self.scenes, self.encodings)):
                for src in e.sources:
                    if src[0] == "clip":
                        c = np.asarray(sc.clips[src[1]][0], np.float32).reshape(-1, 2)
                        first[(i, src[1])] = (n, len(c))
                        pts.append(c)
                        n += len(c)
            arr = np.concatenate(pts) if pts else np.zeros((0, 2), np.float32)
            self._clip_pts = torch.as_tensor(arr, device=dev)
            self._clip_first = first
        return self._clip_pts, self._clip_first

    def update(self, verts, counts, firsts=None):
        """New vertices for the same scenes (the same draws, clips and structure;
        a polygon's vertex count may change): the GPU front end rebuilds the
        frame's lines, path records, bounds, tile bins and plans from them
        (tk_frame_set_polygons) instead of recording and encoding again, and the
        line -> vertex maps come from the lines it kept.  verts: a tensor per scene,
        or one tensor of all scenes' vertices with each scene's first vertex and
        the total in `firsts` (batch_polygons)."""
        dev = torch.device("cuda")
        if firsts is None:
            firsts = np.concatenate([[0], np.cumsum([len(v) for v in verts])]).astype(np.int64)
            verts = torch.cat(list(verts)) if len(verts) else torch.zeros((0, 2), device=dev)
        v32 = verts.detach().to(dev, torch.float32)
        v_total = int(firsts[-1])
        clip_pts, clip_first = self._clip_points(dev)
        key = np.concatenate([np.asarray(firsts, np.int64)] + [np.asarray(k, np.int64) for k in counts])
        hit = getattr(self, "_update_index", None)
        if hit is None or not np.array_equal(hit[0], key) or hit[1].device != dev:
            # each point's row in [scene vertices | clip points], its vertex (-1: a
            # clip point), and the vertex count of each polygon
            rows, p2v, lens = [], [], []
            for i, (e, k) in enumerate(zip(self.encodings, counts)):
                k = np.asarray(k, np.int64)
                offs = firsts[i] + np.concatenate([[0], np.cumsum(k)])
                for src in e.sources:
                    if src[0] == "clip":
                        c0, n = clip_first[(i, src[1])]
                        rows.append(v_total + c0 + np.arange(n))
                        p2v.append(np.full(n, -1, np.int64))
                        lens.append(np.array([n], np.int64))
                    else:
                        lo, hi = src[1], src[2]
                        r = np.arange(offs[lo], offs[hi])
                        rows.append(r)
                        p2v.append(r)
                        lens.append(k[lo:hi])
            cat = lambda parts: np.concatenate(parts) if parts else np.zeros(0, np.int64)
            rows, p2v_np, lens = cat(rows), cat(p2v), cat(lens)
            rows_t = torch.as_tensor(rows, device=dev)
            p2v_t = torch.as_tensor(p2v_np, device=dev)
            lens_t = torch.as_tensor(lens, device=dev)
            starts = torch.zeros(len(lens) + 1, dtype=torch.int64, device=dev)
            starts[1:] = torch.cumsum(lens_t, 0)
            # the point after each one in its polygon
            path = torch.repeat_interleave(torch.arange(len(lens), device=dev), lens_t, output_size=len(rows))
            nxt = torch.arange(len(rows), device=dev) + 1
            nxt = torch.where(nxt == starts[1:][path], starts[:-1][path], nxt)
            hit = (key, rows_t, p2v_t, starts.to(torch.int32), nxt)
            self._update_index = hit
        _, rows_t, p2v, starts32, nxt = hit
        points = torch.cat([v32, clip_pts]).index_select(0, rows_t)
        v_off = v_total
        lv = self.frame.set_polygons(points, starts32, line_vertex=True).long()
        la, lb = p2v[lv], p2v[nxt[lv]]
        # path clip lines: the gradient sink past the vertices
        la = torch.where(la < 0, v_off, la)
        lb = torch.where(lb < 0, v_off, lb)
        self.vertex_count = v_off
        self._la_t, self._lb_t = la, lb
        self.line_a = self.line_b = None  # host copies made on demand (_index)
        self.line_count = len(la)

    def full_paint(self, records):
        """One 32-float paint record per encoded draw (clip draws get zeros) from the
        painter-order path-draw records (DiffScene.paint_records); differentiable."""
        idx = torch.as_tensor(self.draw_index, device=records.device)
        full = torch.zeros((self.draw_count, 32), dtype=records.dtype, device=records.device)
        return full.index_copy(0, idx, records)

    def images(self, flat):
        out = []
        for (w, h, _), off in zip(self.sizes, self.offsets):
            out.append(flat[off:off + w * h].reshape(h, w, 4))
        return out

    def _before(self, device):
        """Orders the context's stream after torch's current stream."""
        if self.sync:
            torch.cuda.synchronize(device)
            return
        if self._stream is None:
            self._stream = torch.cuda.ExternalStream(self.context.stream_handle(), device=device)
        self._stream.wait_stream(torch.cuda.current_stream(device))

    def _after(self, device):
        """Orders torch's current stream after the context's (the outputs, and any
        tensor the call read, are then safe to use or free there)."""
        if self.sync:
            self.context.synchronize()
        else:
            torch.cuda.current_stream(device).wait_stream(self._stream)

    def check_status(self):
        """Waits for the device and raises if the last differentiable render flagged
        an overflow."""
        self.context.synchronize()
        flags = C.c_uint32()
        check(lib().tk_frame_diff_status(self.frame._h, C.byref(flags)))
        if flags.value:
            raise tk.TileinkError(_abi.TK_ERR_INVALID, f"differentiable render flagged {flags.value:#x}")

    def forward(self, colors):
        """colors: (D, 4) premultiplied float32 on the GPU; returns the flat image."""
        colors = colors.detach().to(torch.float32).contiguous()
        image = torch.zeros((self.pixels, 4), dtype=torch.float32, device=colors.device)
        self._before(colors.device)
        check(lib().tk_frame_render_diff(self.frame._h, C.c_void_p(colors.data_ptr()), C.c_void_p(image.data_ptr())))
        self._after(colors.device)
        if self.sync:
            self.check_status()
        self._colors = colors
        return image

    def backward(self, grad_flat):
        grad = grad_flat.detach().to(torch.float32).contiguous()
        dev = grad.device
        line_grad = torch.empty((max(self.line_count, 1), 4), dtype=torch.float32, device=dev)
        color_grad = torch.empty((max(self.draw_count, 1), 32), dtype=torch.float32, device=dev)
        self._before(dev)
        check(lib().tk_frame_backward(self.frame._h, C.c_void_p(self._colors.data_ptr()), C.c_void_p(grad.data_ptr()),
                                      C.c_void_p(line_grad.data_ptr()), C.c_void_p(color_grad.data_ptr())))
        self._after(dev)
        return line_grad[:self.line_count], color_grad[:self.draw_count]

    def vertex_grad(self, line_grad, like):
        la, lb = self._index(like.device)
        g = torch.zeros((self.vertex_count + 1, 2), dtype=like.dtype, device=like.device)
        lg = line_grad.to(like.dtype).to(like.device)
        g.index_add_(0, la, lg[:, 0:2])
        g.index_add_(0, lb, lg[:, 2:4])
        return g[:self.vertex_count]


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, verts, colors, plan, device):
        flat = plan.forward(colors.to(device))
        ctx.plan = plan
        ctx.vdtype, ctx.vdevice = verts.dtype, verts.device
        ctx.cdtype, ctx.cdevice = colors.dtype, colors.device
        return flat

    @staticmethod
    def backward(ctx, grad_flat):
        plan = ctx.plan
        line_grad, color_grad = plan.backward(grad_flat)
        gv = plan.vertex_grad(line_grad, torch.empty(0, dtype=ctx.vdtype, device=ctx.vdevice))
        return gv, color_grad.to(ctx.cdtype).to(ctx.cdevice), None, None


class Renderer:
    """Renders the same scenes again and again as their parameters move (an
    optimiser loop): the first call records and encodes them on the host, later
    calls give the GPU front end only the new vertices (Plan.update).  Calling
    returns the FP32 images, differentiable in the parameters."""

    def __init__(self, context, scenes, device="cuda", sync=True):
        self.context, self.device = context, torch.device(device)
        self.single = isinstance(scenes, DiffScene)
        self.scenes = [scenes] if self.single else list(scenes)
        self.plan = None
        self.sync = sync  # Plan's: False orders the native calls on streams only
        self._cache = {}  # batch_polygons' index tensors

    def __call__(self):
        verts, counts, firsts = batch_polygons(self.scenes, cache=self._cache)
        if self.plan is None:
            self.plan = Plan(self.context, self.scenes, [verts[a:b] for a, b in zip(firsts[:-1], firsts[1:])], counts,
                             sync=self.sync)
        else:
            self.plan.update(verts, counts, firsts)
        flat = _Render.apply(verts, self.plan.full_paint(batch_paint_records(self.scenes)), self.plan, self.device)
        self.flat = flat  # all scenes' pixels in the plan's layout (Plan.images), differentiable
        images = self.plan.images(flat)
        return images[0] if self.single else images


def render(context, scenes, device="cuda"):
    """FP32 images of the scenes (a list of DiffScene), differentiable in their
    parameters.  One scene renders as one frame, several as one batch."""
    single = isinstance(scenes, DiffScene)
    scenes = [scenes] if single else list(scenes)
    verts, counts, firsts = batch_polygons(scenes)
    plan = Plan(context, scenes, [verts[a:b] for a, b in zip(firsts[:-1], firsts[1:])], counts)
    flat = _Render.apply(verts, plan.full_paint(batch_paint_records(scenes)), plan, torch.device(device))
    images = plan.images(flat)
    return images[0] if single else images


def oracle_inputs(scene, verts, counts, recorded=False):
    """The FP64 oracle's arguments for one scene: draws (line_start, line_count, rule)
    and lines (L, 4), from the polygon vertices in their own precision, or, with
    recorded=True, rounded to FP32 as the recorder stores them."""
    v32 = verts.detach().to("cpu", torch.float32).numpy()
    la, lb, per_draw = edge_map(v32, counts)
    v = v32.astype(np.float64) if recorded else verts.detach().cpu().double().numpy()
    lines = np.concatenate([v[la], v[lb]], 1) if len(la) else np.zeros((0, 4))
    starts = np.concatenate([[0], np.cumsum(per_draw)[:-1]]) if len(per_draw) else np.zeros(0, np.int64)
    rules = scene.draw_rules()
    draws = [(int(s), int(n), int(r)) for s, n, r in zip(starts, per_draw, rules)]
    return draws, lines, la, lb


# --- synthetic scenes -----------------------------------------------------------------


def blob_fills(rng, count, width, height, cubics=4, radius_mu=math.log(24.0), radius_sigma=0.5):
    """Closed blobs of `cubics` cubics: ends on a jittered circle, control points a
    third of the way along jittered tangents."""
    cx = rng.uniform(0, width, count)
    cy = rng.uniform(0, height, count)
    r = np.clip(rng.lognormal(radius_mu, radius_sigma, count), 4.0, 128.0)
    ang = np.linspace(0, 2 * math.pi, cubics + 1)[:-1][None, :] + rng.uniform(0, 2 * math.pi, (count, 1))
    rad = r[:, None] * rng.uniform(0.6, 1.2, (count, cubics))
    ex, ey = cx[:, None] + rad * np.cos(ang), cy[:, None] + rad * np.sin(ang)
    ends = np.stack([ex, ey], -1)
    starts = np.roll(ends, 1, axis=1)
    span = ends - starts
    k = 4.0 / 3.0 * math.tan(math.pi / (2 * cubics))
    tang_s = np.stack([-np.sin(np.roll(ang, 1, 1)), np.cos(np.roll(ang, 1, 1))], -1) * np.roll(rad, 1, 1)[..., None]
    tang_e = np.stack([-np.sin(ang), np.cos(ang)], -1) * rad[..., None]
    c1 = starts + k * tang_s + rng.normal(0, 0.15, starts.shape) * np.linalg.norm(span, axis=-1, keepdims=True)
    c2 = ends - k * tang_e + rng.normal(0, 0.15, ends.shape) * np.linalg.norm(span, axis=-1, keepdims=True)
    return np.stack([c1, c2, ends], 2)


def wave_strokes(rng, count, width, height, cubics=2):
    """Open chains of cubics meandering from a random anchor."""
    out = np.zeros((count, 3 * cubics + 1, 2))
    for s in range(count):
        p = np.array([rng.uniform(0, width), rng.uniform(0, height)])
        pts = [p]
        for _ in range(cubics):
            step = rng.normal(0, 18, (3, 2))
            for j in range(3):
                pts.append(pts[-1] + step[j])
        out[s] = np.array(pts)
    return out


def random_gradients(scene, fraction, seed, radial=0.0):
    """Gives a `fraction` of the scene's draws gradient paints of 2-4 stops (a
    `radial` share of them radial, the rest linear): a random line of 16-96 px,
    sorted offsets, opaque-ish stop colours."""
    rng = np.random.default_rng(seed)
    n = len(scene.order)
    for d in np.nonzero(rng.uniform(0, 1, n) < fraction)[0]:
        kind = "radial" if rng.uniform(0, 1) < radial else "linear"
        k = int(rng.integers(2, 5))
        x0, y0 = rng.uniform(0, scene.width), rng.uniform(0, scene.height)
        ang, length = rng.uniform(0, 2 * math.pi), rng.uniform(16, 96)
        line = (x0, y0, x0 + length * math.cos(ang), y0 + length * math.sin(ang))
        offsets = np.sort(rng.uniform(0, 1, k))
        offsets[0], offsets[-1] = 0.0, 1.0
        stops = np.concatenate([rng.uniform(0, 1, (k, 3)), rng.uniform(0.3, 1.0, (k, 1))], 1)
        scene.gradients[int(d)] = (line, offsets, torch.tensor(stops, dtype=scene.colors.dtype,
                                                              device=scene.colors.device), kind)
    return scene


def random_scene(seed, width, height, fills, strokes=0, cubics=4, dtype=torch.float64, device="cpu",
                 radius_mu=math.log(24.0), stroke_cubics=2):
    rng = np.random.default_rng(seed)
    f = blob_fills(rng, fills, width, height, cubics, radius_mu)
    s = wave_strokes(rng, strokes, width, height, stroke_cubics) if strokes else None
    widths = rng.lognormal(math.log(3.0), 0.4, strokes) if strokes else None
    colors = np.concatenate([rng.uniform(0, 1, (fills + strokes, 3)), rng.uniform(0.3, 0.9, (fills + strokes, 1))], 1)
    order = [(0, i) for i in range(fills)] + [(1, i) for i in range(strokes)]
    rng.shuffle(order)
    t = lambda a: None if a is None else torch.tensor(a, dtype=dtype, device=device)
    return DiffScene(width, height, fills=t(f), strokes=t(s), widths=t(widths), colors=t(colors), order=order,
                     rules=rng.integers(0, 2, fills))
