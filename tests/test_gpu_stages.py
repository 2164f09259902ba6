"""Stage entry points of the C ABI on the GPU, replaying the reference's known-answer
cases verbatim (guard words included) and comparing every other output with the
oracle bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import _abi, scenes
from tileink_amd._abi import check, lib

from test_oracle_kats import (LERP_CASES, PREFIX_LENGTHS, PREFIX_SEEDS, _coverage_requests, count_case_buffers,
                              count_segments, prefix_case)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lengths", PREFIX_LENGTHS)
@pytest.mark.parametrize("seed", PREFIX_SEEDS)
def test_scan_prefix_kat_single_pass(ctx, lengths, seed):
    inp, exp = prefix_case(lengths, seed)
    n = len(inp["counts"])
    cfg = _abi.ScanConfig(*inp["config"])
    d_chunks = ctx.to_device(inp["chunks"])
    d_ranges = ctx.to_device(np.full(2 * n, 0x5555_5555, np.uint32))
    d_cursors = ctx.to_device(np.full(n, 0x5555_5555, np.uint32))
    d_counts = ctx.to_device(inp["counts"])
    d_totals = ctx.to_device(np.zeros(len(lengths), np.uint32))
    d_offsets = ctx.to_device(np.zeros(len(lengths), np.uint32))
    d_bumps = ctx.to_device(np.full(2, 0x5555_5555, np.uint32))
    d_paths = ctx.to_device(inp["paths"])
    d_cr = ctx.to_device(inp["chunk_ranges"])
    check(lib().tk_scan_prefix(ctx.handle, C.byref(cfg), d_chunks.ptr, d_cr.ptr, d_paths.ptr, d_counts.ptr,
                               d_ranges.ptr, d_cursors.ptr, d_bumps.ptr, d_totals.ptr, d_offsets.ptr))
    ctx.synchronize()
    assert np.array_equal(d_ranges.read(np.uint32, 2 * n), exp["ranges"])
    assert np.array_equal(d_cursors.read(np.uint32, n), exp["cursors"])
    assert np.array_equal(d_totals.read(np.uint32, len(lengths)), exp["totals"])
    assert np.array_equal(d_offsets.read(np.uint32, len(lengths)), exp["offsets"])
    assert np.array_equal(d_bumps.read(np.uint32, 2), exp["bumps"])
    assert np.array_equal(d_counts.read(np.uint32, n), exp["counts"])


def test_scan_count_kat_matches_cpu_counts_and_keeps_guards(ctx):
    for p0, p1 in count_segments():
        cfg, lines, path, init = count_case_buffers(p0, p1)
        d_lines, d_path = ctx.to_device(lines), ctx.to_device(path)
        d_bd, d_counts = ctx.to_device(init), ctx.to_device(init)
        check(lib().tk_scan_count(ctx.handle, C.byref(_abi.ScanConfig(*cfg)), d_lines.ptr, d_path.ptr, 1,
                                  d_bd.ptr, d_counts.ptr))
        ctx.synchronize()
        bd, counts = d_bd.read(np.uint32, 78), d_counts.read(np.uint32, 78)
        for buf in (bd, counts):
            assert np.all(buf[:7] == 0x1234_5678) and np.all(buf[71:] == 0x1234_5678), (p0, p1)
        assert int(counts[7:71].sum()) == tk.line_scanned_tile_count(p0, p1, (0, 0, 8, 8)), (p0, p1)
        # and every word agrees with the oracle (backdrop deltas included)
        ob, oc = init.copy().view(np.int32), init.copy()
        oracle.scan_count(oracle.ScanConfig(*cfg), lines, path, 1, ob, oc)
        assert np.array_equal(bd, ob.view(np.uint32)) and np.array_equal(counts, oc), (p0, p1)


@pytest.mark.parametrize("entry", ["count", "emit"])
def test_invalid_path_references_leave_outputs_untouched(ctx, entry):
    # scan_invalid_path_references_leave_count_and_emit_outputs_untouched (metal/scan_tests.rs:97-141)
    cfg = _abi.ScanConfig(0, 0, 1, 0, 3, 8, 0, 0, 0, 0, 0)
    lines = np.array([[p, 0, 0, 0, np.float32(16).view(np.uint32), np.float32(16).view(np.uint32)]
                      for p in (1, 0xFFFFFFFF // 19, 0xFFFFFFFF)], np.uint32)
    d_lines = ctx.to_device(lines)
    d_paths = ctx.to_device(np.zeros(19, np.uint32))
    guard = np.full(128, 0xa5, np.uint8)
    first, second = ctx.to_device(guard), ctx.to_device(guard)
    if entry == "count":
        check(lib().tk_scan_count(ctx.handle, C.byref(cfg), d_lines.ptr, d_paths.ptr, 1, first.ptr, second.ptr))
    else:
        check(lib().tk_scan_emit(ctx.handle, C.byref(cfg), d_lines.ptr, d_paths.ptr, 1, first.ptr, second.ptr))
    ctx.synchronize()
    assert np.array_equal(first.read(np.uint8, 128), guard)
    assert np.array_equal(second.read(np.uint8, 128), guard)


def test_pixel_math_words_match_oracle(ctx):
    rng = np.random.default_rng(11)
    n = 20000
    req = np.zeros((n, 8), np.uint32)
    req[:, 0] = rng.integers(0, 256, n)
    req[:, 1] = rng.integers(0, 256, n)
    req[:, 2] = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    req[:, 3] = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    req[:, 4] = rng.uniform(-3, 3, n).astype(np.float32).view(np.uint32)
    req[:, 5] = rng.uniform(0, 1, n).astype(np.float32).view(np.uint32)
    req[:, 6] = rng.integers(0, 2, n)
    for i, (a, b, t, _) in enumerate(LERP_CASES):
        req[i] = [0, 0, a * 0x0101_0101, b * 0x0101_0101, 0, t, 0, 0]
    d_req = ctx.to_device(req)
    d_out = ctx.to_device(np.zeros((n, 14), np.uint32))
    check(lib().tk_pixel_math_words(ctx.handle, n, d_req.ptr, d_out.ptr))
    ctx.synchronize()
    got = d_out.read(np.uint32, n * 14).reshape(n, 14)
    assert np.array_equal(got, oracle.pixel_math_words(req))
    assert got[:4, 7].tolist() == [e * 0x0101_0101 for *_, e in LERP_CASES]


def test_fill_coverage_fp32_matches_oracle_and_fp64_area(ctx):
    a = scenes.mixed_scene_640x360().arrays()
    _, out = oracle.render_frame(a)
    req = _coverage_requests(a, out, limit=60000, seed=2)
    segs = out["segments"]
    d_req, d_segs = ctx.to_device(req), ctx.to_device(segs)
    n = len(req)
    d_alpha = ctx.to_device(np.zeros(n, np.uint32))
    d_cov = ctx.to_device(np.zeros(n, np.float32))
    check(lib().tk_fill_coverage_words(ctx.handle, n, d_req.ptr, d_segs.ptr, d_alpha.ptr, d_cov.ptr))
    ctx.synchronize()
    alpha, cov = d_alpha.read(np.uint32, n), d_cov.read(np.float32, n)
    o_alpha, o_cov = oracle.fill_coverage_words(req, segs)
    assert np.array_equal(alpha, o_alpha)
    assert np.array_equal(cov.view(np.uint32), o_cov.view(np.uint32))
    exact = oracle.fill_coverage_f64(req, segs)
    assert np.abs(cov.astype(np.float64) - exact).max() <= 1e-5


def test_cumsum_rows_match_numpy(ctx):
    rng = np.random.default_rng(4)
    row_lens = rng.integers(1, 700, 300)
    offs, lens, starts, ends, pos = [], [], [], [], 0
    for n in row_lens:
        starts.append(len(offs))
        for k in range(0, n, 256):
            offs.append(pos + k)
            lens.append(min(256, n - k))
        ends.append(len(offs))
        pos += n
    vals = rng.integers(-3, 4, pos).astype(np.int32)
    expected = np.concatenate([np.cumsum(vals[s:s + n]) for s, n in zip(np.cumsum(row_lens) - row_lens, row_lens)])
    d_vals = ctx.to_device(vals)
    arrs = [ctx.to_device(np.array(x, np.uint32)) for x in (offs, lens, starts, ends)]
    d_tot, d_car = ctx.to_device(np.zeros(len(offs), np.uint32)), ctx.to_device(np.zeros(len(offs), np.uint32))
    cfg = _abi.CumsumConfig(len(starts), len(offs), 0, 0)
    check(lib().tk_cumsum(ctx.handle, C.byref(cfg), *[b.ptr for b in arrs], d_vals.ptr, d_tot.ptr, d_car.ptr))
    ctx.synchronize()
    assert np.array_equal(d_vals.read(np.int32, pos), expected.astype(np.int32))


def test_render_to_image_entry_matches_frame(ctx):
    canvas = scenes.clipped_paths_small(32)
    renderer = tk.NativeRenderer(tk.NativeBackend.Hip, 128, 128, context=ctx)
    image = renderer.render_to_image(canvas).readback()
    ref, _ = oracle.render_frame(canvas.arrays())
    assert np.array_equal(image, ref)
