"""The reference's coarse known-answer scenes, restated as data, and two runners:
the oracle's per-batch coarse_count_bins / coarse_emit_bins (fed the reference's
own precomputed tile record, as its tests feed the shader), and the device's
tk_coarse (count, prefix and emit of every batch in one call).

Scenes (src/native/runtime/tests/coarse_count/):
- count_scene (scene.rs:7-106): one tile, a list of `draws` entries cycling
  over 8 draw kinds (SDF, SDF shadow, fill, path glyph, opacity-tagged, other
  batch, clip-tagged, no geometry), optionally relocated behind 13 unrelated
  indices, under stack mode 0 (none), 1 (full path clip, elided), 2 (opacity,
  retained) or 3 (empty path clip, tile rejected);
- emission_scene (emit_scene.rs:3-62): the stream count_scene emits, with a
  capacity guard;
- paint_scene (emit_scene.rs:64-200): one draw whose paint picks each fast
  path (COLOR, FILL, SDF, IMAGE, PATH_GLYPH) and tile kind;
- the nested END-order stream (emit.rs:62-90);
- mixed_stack (stack.rs:5-101): full, partial, rejected and even-odd clips
  with opacity and blend wrappers;
- opaque_scene (occlusion.rs:6-79): the opaque-prefix skip word.
Records are PtclRecord words {tag, backdrop, fill_rule, segment_start,
segment_end, color}; kinds are tile kind | skip << 3.
"""
import ctypes as C

import numpy as np

import oracle

NONE = 0xFFFFFFFF
GUARD = 0x4545_4545


def f32(v):
    return int(np.float32(v).view(np.uint32))


class Kat:
    """One coarse test scene with its expectations.

    tile_record: the reference's precomputed TileCoarseRecord for the emit KATs;
    stream: the records emit writes (END included; empty when nothing is
    emitted); count: what count_bins counts; kind: the tile kind word."""

    def __init__(self):
        self.paint = np.zeros(9, np.uint32)
        self.brush_base = 0
        self.batch = 42
        self.layer_range = (0, 0)
        self.capacity = 3
        self.tile_record = None
        self.stream = []
        self.count = 0
        self.kind = 1

    # -- oracle -------------------------------------------------------------
    def oracle_config(self):
        return oracle.CoarseConfig(tile_count=1, tiles_width=1, tiles_height=1, batch_id=self.batch,
                                   layer_stack_start=self.layer_range[0], layer_stack_end=self.layer_range[1],
                                   ptcl_capacity=self.capacity, paint_brush_base=self.brush_base,
                                   draw_count=len(self.draws), path_count=len(self.paths),
                                   paint_words=len(self.paint), layer_count=len(self.layers),
                                   tile_draw_index_count=len(self.indices))

    def _oracle_inputs(self):
        keep = [np.ascontiguousarray(a, dtype=np.uint32) for a in
                (self.draws, self.paths, self.backdrops, self.ranges, self.layers, self.paint, self.batch_of,
                 np.array([self.list_start, self.list_len], np.uint32), self.indices)]
        return oracle.CoarseInputs(*[k.ctypes.data for k in keep]), keep

    def oracle_count(self):
        inputs, keep = self._oracle_inputs()
        work = np.full(6 + 6 * self.capacity, GUARD, np.uint32)
        kinds = np.full(1, GUARD, np.uint32)
        oracle.coarse_count_bins(self.oracle_config(), inputs, work, kinds)
        return work, kinds

    def oracle_emit(self):
        """Runs emit on the reference's tile record; returns (PTCL slots, kind)."""
        inputs, keep = self._oracle_inputs()
        work = np.full(6 + 6 * self.capacity, GUARD, np.uint32)
        work[:6] = self.tile_record
        kinds = np.full(1, GUARD, np.uint32)
        oracle.coarse_emit_bins(self.oracle_config(), inputs, work, kinds)
        return work[6:].reshape(-1, 6), int(kinds[0])

    def expected_slots(self, stream=None, skip_end=False):
        """The PTCL area after emit: the stream from the tile record's start,
        records at or past the capacity dropped, the rest untouched."""
        slots = np.full((self.capacity, 6), GUARD, np.uint32)
        start = int(self.tile_record[1])
        for i, rec in enumerate(self.stream if stream is None else stream):
            if start + i < self.capacity:
                slots[start + i] = rec
        return slots

    # -- device -------------------------------------------------------------
    def device(self, ctx, capacity=None):
        """tk_coarse on the scene: its batch is batch 0 of the call, draws of any
        other batch go to batch 1 (which has no layers).  Returns (count, stream
        records of batch 0 in [start, end) within the written capacity, kind,
        status words)."""
        import tileink_amd as tk
        from tileink_amd import _abi
        from tileink_amd._abi import check, lib
        cap = self.capacity if capacity is None else capacity
        batch_of = np.full(len(self.draws), NONE, np.uint32)
        batch_of[:min(len(self.batch_of), len(self.draws))] = self.batch_of[:len(self.draws)]
        batch_ids = np.where(batch_of == self.batch, 0, 1).astype(np.uint32)
        batches = np.array([[self.layer_range[0], self.layer_range[1]], [0, 0]], np.uint32)
        bufs = {
            "draws": self.draws, "paths": self.paths, "backdrops": self.backdrops, "segment_ranges": self.ranges,
            "layers": self.layers, "batches": batches, "paint": self.paint, "draw_batch_ids": batch_ids,
            "tile_draws": np.array([self.list_start, self.list_len], np.uint32), "tile_draw_indices": self.indices,
        }
        dev = {k: ctx.to_device(np.ascontiguousarray(v, dtype=np.uint32).reshape(-1)) for k, v in bufs.items()}
        Q = 2
        counts = ctx.to_device(np.full(Q, GUARD, np.uint32))
        ranges = ctx.to_device(np.full(2 * Q, GUARD, np.uint32))
        kinds = ctx.to_device(np.full(Q, GUARD, np.uint32))
        bases = ctx.to_device(np.zeros(Q + 1, np.uint32))
        ptcl = ctx.to_device(np.full(6 * max(cap, 1), GUARD, np.uint32))
        status = ctx.to_device(np.zeros(4, np.uint32))
        cfg = _abi.CoarseConfig(tile_count=1, tiles_width=1, tiles_height=1, batch_count=Q,
                                draw_count=len(self.draws), path_count=len(self.paths),
                                paint_words=len(self.paint), paint_brush_base=self.brush_base,
                                layer_count=len(self.layers), tile_draw_index_count=len(self.indices),
                                ptcl_capacity=cap)
        ins = _abi.CoarseInputs(*[dev[k].ptr for k in ("draws", "paths", "backdrops", "segment_ranges", "layers",
                                                        "batches", "paint", "draw_batch_ids", "tile_draws",
                                                        "tile_draw_indices")])
        outs = _abi.CoarseOutputs(counts.ptr, ranges.ptr, kinds.ptr, bases.ptr, ptcl.ptr, status.ptr)
        check(lib().tk_coarse(ctx.handle, C.byref(cfg), C.byref(ins), C.byref(outs)))
        ctx.synchronize()
        n = int(counts.read(np.uint32, Q)[0])
        start, end = (int(v) for v in ranges.read(np.uint32, 2 * Q)[:2])
        words = ptcl.read(np.uint32, 6 * max(cap, 1)).reshape(-1, 6)
        stream = words[start:min(end, cap)]
        total = int(bases.read(np.uint32, Q + 1)[Q])
        return n, stream, int(kinds.read(np.uint32, Q)[0]), status.read(np.uint32, 4), (start, end, total)


def count_scene(draw_count, relocated, stack):
    """scene.rs:7-106."""
    pattern = [1, 1, 1, 1, 0, 0, 1, 0]
    draws = np.zeros((8, 31), np.uint32)
    for i in range(8):
        d = draws[i]
        d[0], d[1], d[2], d[4], d[12], d[13] = 0, NONE, NONE, NONE, 16, 16
        d[19] = d[22] = f32(1.0)
        if i == 0:
            d[2] = 0
        elif i == 1:
            d[4] = 0
        elif i == 3:
            d[8] = 5
        elif i == 4:
            d[8] = 2
        elif i == 6:
            d[8] = 1
        elif i == 7:
            d[0] = NONE
    k = Kat()
    prefix = 13 if relocated else 0
    k.indices = np.full(prefix + draw_count, GUARD, np.uint32)
    k.indices[prefix:] = np.arange(draw_count) % 8
    k.list_start = NONE if draw_count == 0 else prefix
    k.list_len = draw_count
    particles = sum(pattern[i % 8] for i in range(draw_count))
    wrapper = 1 if stack == 2 else 0
    if stack == 3:
        particles = 0
    k.count = 0 if particles == 0 else particles + 2 * wrapper + 1
    k.kind = 1 if k.count == 0 else 0
    k.draws = draws
    path = np.zeros((1, 19), np.uint32)
    path[0, 8] = path[0, 9] = 1
    k.paths = path
    k.backdrops = np.array([1 if stack != 3 else 0], np.uint32)
    k.ranges = np.zeros(2, np.uint32)
    k.layers = np.array([[0 if stack in (1, 3) else (1 if stack == 2 else 99), 2, 0]], np.uint32)
    k.layer_range = (0, 1 if stack != 0 else 0)
    k.batch_of = np.array([42, 42, 42, 42, 42, 99, 42, 42], np.uint32)
    k.capacity = 3
    return k


def emission_scene(draws, relocated, stack, capacity):
    """emit_scene.rs:3-62 with the adjustments of emit.rs:5-34 for the bins emitter."""
    k = count_scene(draws, relocated, stack)
    k.capacity = capacity
    k.draws[:, 6] = NONE
    stream = []
    for index in range(draws):
        d = index % 8
        if d in (0, 1):
            stream.append([9, 0, 0, d, 0, d])
        elif d == 2:
            stream.append([1, 1, 0, 0, 0, 2])
        elif d == 3:
            stream.append([11, 1, 0, 0, 0, 3])
        elif d == 6:
            stream.append([3, 1, 0, 0, 0, 0])
    nonempty = bool(stream)
    if nonempty and stack == 2:
        stream = [[5, 1, 0, 0, 0, 0]] + stream + [[6, 0, 0, 0, 0, 0]]
    if nonempty:
        stream.append([0] * 6)
    k.tile_record = [len(stream), 1, 1 + len(stream), 0, 1, 1]
    if stack == 3:
        # The stack rejects the tile: emit writes one END at the record's start.
        k.stream = [[0] * 6] if nonempty else []
        k.device_stream = []
        k.kind = 1
    else:
        k.stream = stream
        k.device_stream = stream
        k.kind = 0 if nonempty else 1
    return k


PAINT_MODES = [m for m in range(21) if m not in (14, 15, 16)]
"""paint_scene modes; 14-16 are glyph (text) particles, outside this path."""


def paint_scene(mode):
    """emit_scene.rs:64-200: returns the scene with stream [record, END]."""
    k = count_scene(1, False, 0)
    k.capacity = 8
    k.draws[:, 6] = NONE
    d = k.draws[0]
    d[6] = 0
    d[2] = 0
    d[3] = 9
    d[18] = 1
    k.brush_base = 12
    p = np.zeros(20, np.uint32)
    p[0] = 1
    p[3] = p[4] = f32(16.0)
    p[12] = 1
    p[16] = 0xFF33_6699
    record = [9, 0, 0, 0, 0, 0]
    kind = 3
    if mode == 0:
        d[2] = NONE
        record, kind = [2, 1, 0, 0, 0, 0xFF33_6699], 2
    elif mode == 1:
        d[2] = NONE
        d[20] = f32(0.125)
        record, kind = [1, 1, 0, 0, 0, 0], 0
    elif mode == 2:
        d[2] = NONE
        p[16] = 0
        record, kind = [1, 1, 0, 0, 0, 0], 0
    elif mode == 3:
        record, kind = [2, 0, 0, 0, 0, 0xFF33_6699], 2
    elif mode == 4:
        p[5] = f32(4.0)
    elif mode == 5:
        d[19] = f32(2.0)
    elif mode == 6:
        d[4] = 0
        kind = 0
    elif mode in (7, 8):
        p[0] = 5 if mode == 7 else 14
        d[3] = 1
    elif mode == 9:
        p[0] = 2
        kind = 0
    elif mode in (10, 11, 17):
        p[12] = 7
        p[19] = 254 if mode == 11 else 255
        if mode == 11:
            kind = 0
        else:
            record = [13, 0, 0, 0, 0, 0]
        if mode == 17:
            p[1] = p[2] = f32(16.0)
            p[3] = p[4] = 0
    elif mode == 12:
        d[6] = NONE
        kind = 0
    elif mode == 13:
        p[16] = 0
        kind = 0
    elif mode == 18:
        d[2] = NONE
        d[4] = 0
        kind = 0
    elif mode == 19:
        d[2] = NONE
        d[8] = 5
        record, kind = [11, 1, 0, 0, 0, 0], 0
    elif mode == 20:
        d[2] = NONE
        d[6] = 2
        p[12] = 0
        p[14] = 1
        p[18] = 0xFF33_6699
        k.backdrops[0] = NONE
        record, kind = [2, NONE, 0, 0, 0, 0xFF33_6699], 2
    else:
        raise ValueError(mode)
    k.paint = p
    k.tile_record = [2, 1, 3, 0, 1, 4]
    k.stream = k.device_stream = [record, [0] * 6]
    k.count = 2
    k.kind = kind
    return k


def nested_end_order_scene():
    """emit.rs:62-90: opacity, a full (elided) path clip, blend and an SDF clip
    wrap paint mode 4's SDF particle; END records come back in reverse."""
    k = paint_scene(4)
    k.capacity = 16
    k.layer_range = (1, 5)
    k.layers = np.array([[99, 99, 99], [1, 2, 0x1234], [0, 6, 0], [2, 3, 17], [0, 0, 0]], np.uint32)
    k.stream = k.device_stream = [[5, 1, 0, 0, 0, 0x1234], [7, 1, 0, 0, 0, 17], [12, 0, 0, 0, 0, 0],
                                  [9, 0, 0, 0, 0, 0], [4, 0, 0, 0, 0, 0], [8, 0, 0, 0, 0, 0], [6, 0, 0, 0, 0, 0],
                                  [0] * 6]
    k.tile_record = [8, 1, 9, 0, 1, 1]
    k.count = 8
    k.kind = 0
    return k


def mixed_stack(case):
    """stack.rs:5-101 (bins emitter): clip path draw 6 (full; case 1 partial;
    case 4 outside; case 5 even-odd under winding -1), opacity draw 2, blend
    draw 3 and an SDF clip draw 8 (radius 4; case 2 full; case 3 outside)."""
    k = paint_scene(4)
    k.capacity = 32
    clip = k.draws[0].copy()
    clip[0] = NONE
    clip[2] = len(k.paint)
    clip[3] = 9
    clip[6] = NONE
    clip[8] = 1
    clip[18] = 0
    sdf = [1, 0, 0, f32(16.0), f32(16.0), 0 if case == 2 else f32(4.0), 0, 0, 0]
    if case == 3:
        clip[12:14] = 0
        sdf[3] = sdf[4] = 0
    k.paint = np.concatenate([k.paint, np.array(sdf, np.uint32)])
    k.draws = np.concatenate([k.draws, clip[None]])
    k.draws[2, 8] = 2
    k.draws[3, 8] = 3
    k.layer_range = (1, 5)
    k.layers = np.array([[99, 99, 99], [0, 6, 0], [1, 2, 128], [2, 3, 17], [0, 8, 0]], np.uint32)
    if case == 1:
        k.ranges[1] = 1
    if case == 4:
        k.paths[0, 8] = 0
        k.draws[6, 12:14] = 0
    if case == 5:
        k.backdrops[0] = NONE
        k.draws[6, 9] = 1
    rejected = case in (3, 4)
    winding, end = int(k.backdrops[0]), int(k.ranges[1])
    stream = []
    if case == 1:
        stream.append([3, winding, 0, 0, end, 0])
    stream += [[5, winding, 0, 0, end, 128], [7, winding, 0, 0, end, 17]]
    if case != 2:
        stream.append([12, 0, 0, 0, 0, 8])
    stream.append([9, 0, 0, 0, 0, 0])
    if case != 2:
        stream.append([4, 0, 0, 0, 0, 0])
    stream += [[8, 0, 0, 0, 0, 0], [6, 0, 0, 0, 0, 0]]
    if case == 1:
        stream.append([4, 0, 0, 0, 0, 0])
    stream.append([0] * 6)
    k.tile_record = [len(stream), 1, 1 + len(stream), 0, 1, 1]
    k.count = 0 if rejected else len(stream)
    if rejected:
        k.stream = [[0] * 6]
        k.device_stream = []
        k.kind = 1
    else:
        k.stream = k.device_stream = stream
        k.kind = 0
    return k


def opaque_scene(count, relocated, stack, alpha):
    """occlusion.rs:6-79 (bins emitter): `count` full-tile solid SDF draws of
    alpha 119 except the middle one (`alpha`); with alpha 255 and no wrapper the
    tile kind records the records before it as a skippable prefix."""
    k = count_scene(count, relocated, stack)
    k.capacity = count + 4
    paint = paint_scene(3)
    k.brush_base = 12
    sdf = [paint.paint[:12]]
    group = k.draws[2].copy()
    n = max(count, 3)
    hidden = count // 2
    draws = []
    for index in range(n):
        d = paint.draws[0].copy()
        d[6] = index * 21
        draws.append(d)
        brush = np.zeros(21, np.uint32)
        brush[0] = 1
        brush[4] = ((alpha if index == hidden else 119) << 24) | 0x0022_4466
        sdf.append(brush)
    k.batch_of = np.full(n + (1 if stack else 0), 42, np.uint32)
    if stack != 0:
        group[8] = 1 if stack == 1 else 2
        k.layers[0, 1] = n
        draws.append(group)
    k.draws = np.array(draws, np.uint32)
    k.paint = np.concatenate(sdf).astype(np.uint32)
    k.indices = np.full((13 if relocated else 0) + count, GUARD, np.uint32)
    k.indices[(13 if relocated else 0):] = np.arange(count)
    wrappers = 1 if stack == 2 else 0
    length = 0 if count == 0 else count + 2 * wrappers + 1
    k.tile_record = [length, 1, 1 + length, 0, 1, 1]
    stream = []
    if count:
        if wrappers:
            stream.append([5, 1, 0, 0, 0, 0])
        for index in range(count):
            stream.append([2, 0, 0, 0, 0, ((alpha if index == hidden else 119) << 24) | 0x0022_4466])
        if wrappers:
            stream.append([6, 0, 0, 0, 0, 0])
        stream.append([0] * 6)
    k.stream = k.device_stream = stream
    k.count = length
    k.kind = 1 if count == 0 else (0 if wrappers else 2)
    if count > 1 and alpha == 255 and not wrappers:
        k.kind |= hidden << 3
    return k
