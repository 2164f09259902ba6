"""Shared parity checks: the gfx950 frame against the CPU oracle on one scene."""
import numpy as np

import oracle
import tileink_amd as tk
from tileink_amd import _abi


def gpu_frame(ctx, canvas, clear=0):
    scene = canvas.encode(clear)
    frame = tk.Frame(ctx, scene)
    frame.set_capture(True)
    frame.render()
    return frame, scene


def read_frame(frame, scene):
    stats = frame.stats()
    B, P, S = scene.backdrop_len, scene.path_count, scene.segment_capacity
    QT = scene.batch_count * scene.tiles_width * scene.tiles_height
    Q = scene.batch_count
    out = {
        "image": frame.image(),
        "raw_backdrops": frame.buffer(_abi.BUF_RAW_BACKDROPS, np.int32, B),
        "backdrops": frame.buffer(_abi.BUF_BACKDROPS, np.int32, B),
        "segment_counts": frame.buffer(_abi.BUF_SEGMENT_COUNTS, np.uint32, B),
        "segment_ranges": frame.buffer(_abi.BUF_SEGMENT_RANGES, np.uint32, 2 * B).reshape(-1, 2),
        "segment_cursors": frame.buffer(_abi.BUF_SEGMENT_CURSORS, np.uint32, B),
        "segment_bumps": frame.buffer(_abi.BUF_SEGMENT_BUMPS, np.uint32, P),
        "segments": frame.buffer(_abi.BUF_SEGMENTS, np.uint32, 5 * S).reshape(-1, 5),
        "tile_ptcl_counts": frame.buffer(_abi.BUF_TILE_PTCL_COUNTS, np.uint32, QT),
        "tile_ptcl_ranges": frame.buffer(_abi.BUF_TILE_PTCL_RANGES, np.uint32, 2 * QT).reshape(-1, 2),
        "tile_kinds": frame.buffer(_abi.BUF_TILE_KINDS, np.uint32, QT),
        "batch_bases": frame.buffer(_abi.BUF_BATCH_PTCL_BASES, np.uint32, Q + 1),
        "stats": stats,
    }
    out["ptcl"] = frame.buffer(_abi.BUF_PTCL, np.uint32, 6 * int(stats["ptcl_records"])).reshape(-1, 6)
    return out


def compare_frame(ctx, canvas, clear=0, hashed_segments=False):
    """Returns (gpu, ref, mismatches) where mismatches lists the differing stages.
    hashed_segments: compare each tile's segment multiset through an order-free
    hash sum instead of sorting (for frames with tens of millions of segments)."""
    frame, scene = gpu_frame(ctx, canvas, clear)
    gpu = read_frame(frame, scene)
    del frame
    arrays = canvas.arrays(clear)
    ref_image, ref = oracle.render_frame(arrays)
    ref["image"] = ref_image
    bad = []
    # The device reads the recorder's tile bins; the oracle derives its own.
    tile_draws, indices = oracle.tile_draw_bins(arrays)
    if not (np.array_equal(arrays.tile_draws, tile_draws) and np.array_equal(arrays.tile_draw_indices, indices)):
        bad.append("tile_bins")
    for key in ("raw_backdrops", "backdrops", "segment_counts", "segment_bumps", "segment_cursors",
                "tile_ptcl_counts", "tile_kinds", "image"):
        if not np.array_equal(gpu[key], ref[key]):
            bad.append(key)
    if not np.array_equal(gpu["segment_ranges"], ref["segment_ranges"]):
        bad.append("segment_ranges")
    elif hashed_segments:
        if not np.array_equal(tile_segment_hashes(gpu["segments"], gpu["segment_ranges"]),
                              tile_segment_hashes(ref["segments"], ref["segment_ranges"])):
            bad.append("segments")
    elif not np.array_equal(tile_sorted_segments(gpu["segments"], gpu["segment_ranges"]),
                            tile_sorted_segments(ref["segments"], ref["segment_ranges"])):
        bad.append("segments")
    if not np.array_equal(gpu["batch_bases"], ref["batch_bases"]):
        bad.append("batch_bases")
    T = scene.tiles_width * scene.tiles_height
    if scene.batch_count:
        bases = np.repeat(gpu["batch_bases"][:-1], T)
        starts = gpu["tile_ptcl_ranges"][:, 0] - bases
        ends = gpu["tile_ptcl_ranges"][:, 1] - bases
        if not (np.array_equal(starts, ref["tile_ptcl_starts"]) and np.array_equal(ends, ref["tile_ptcl_ends"])):
            bad.append("tile_ptcl_ranges")
    if not np.array_equal(gpu["ptcl"], ref["ptcl"]):
        bad.append("ptcl")
    if gpu["stats"]["overflow"]:
        bad.append("overflow")
    return gpu, ref, bad


def tile_segment_hashes(segments, ranges):
    """Per backdrop tile, the wrapping sum of a 64-bit hash of each of its segments'
    five words: equal for equal per-tile segment multisets, in any slot order."""
    ranges = np.asarray(ranges, np.int64).reshape(-1, 2)
    words = np.asarray(segments, np.uint32).reshape(-1, 5)
    h = np.zeros(len(words), np.uint64)
    with np.errstate(over="ignore"):
        for k in range(5):
            h = (h ^ words[:, k].astype(np.uint64)) * np.uint64(0x100000001B3)
            h ^= h >> np.uint64(29)
        h *= np.uint64(0x9E3779B97F4A7C15)
        csum = np.concatenate([np.zeros(1, np.uint64), np.cumsum(h, dtype=np.uint64)])
        end = np.clip(ranges[:, 1], 0, len(words))
        start = np.clip(ranges[:, 0], 0, end)
        return csum[end] - csum[start]


def tile_sorted_segments(segments, ranges):
    """Each backdrop tile's segments sorted by their five words, tiles in order: the
    segment multiset per tile, independent of emit's atomic slot order."""
    ranges = np.asarray(ranges, np.int64).reshape(-1, 2)
    counts = np.clip(ranges[:, 1] - ranges[:, 0], 0, None)
    total = int(counts.sum())
    tile = np.repeat(np.arange(len(ranges)), counts)
    first = np.repeat(ranges[:, 0], counts)
    offset = np.arange(total) - np.repeat(np.cumsum(counts) - counts, counts)
    rows = np.asarray(segments, np.uint32).reshape(-1, 5)[first + offset]
    order = np.lexsort((rows[:, 4], rows[:, 3], rows[:, 2], rows[:, 1], rows[:, 0], tile))
    return np.concatenate([tile[order, None].astype(np.uint32), rows[order]], axis=1)


def channel_diff(a, b):
    return int(np.abs(a.view(np.uint8).astype(np.int32) - b.view(np.uint8).astype(np.int32)).max(initial=0))
