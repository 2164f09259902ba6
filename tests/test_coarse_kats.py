"""The reference's coarse count / emit known-answer scenes on the oracle
(src/native/runtime/tests/coarse_count/{scene,emit,emit_scene,stack,occlusion}.rs),
with the reference's own precomputed tile records, capacities and guard words.
The scenes live in coarse_kats.py; the device runs them in test_gpu_coarse_kats.py.

Not restated: paint modes 14-16 and the glyph-page scenes (glyph particles: text
is outside this path), and the tests of the chunked emit route
(coarse_emit_chunk_*, FrameOptions.chunked, default false), which this path does
not build.  The opaque-prefix rules those chunk tests probe are the ones
opaque_scene checks through the bins emitter."""
import numpy as np
import pytest

from coarse_kats import (GUARD, PAINT_MODES, count_scene, emission_scene, mixed_stack, nested_end_order_scene,
                         opaque_scene, paint_scene)

DRAWS = [0, 1, 255, 256, 257, 513]


def check_emit(k):
    slots, kind = k.oracle_emit()
    assert np.array_equal(slots, k.expected_slots()), (slots[:4].tolist(), k.stream[:4])
    assert kind == k.kind


@pytest.mark.parametrize("draws", DRAWS)
@pytest.mark.parametrize("relocated", [False, True])
@pytest.mark.parametrize("stack", range(4))
def test_count_scene(draws, relocated, stack):
    k = count_scene(draws, relocated, stack)
    work, kinds = k.oracle_count()
    assert work[0] == k.count and work[3] == 0
    assert kinds[0] == (1 if k.count == 0 else 0)
    assert np.all(work[6:] == GUARD)


@pytest.mark.parametrize("draws", DRAWS)
@pytest.mark.parametrize("relocated", [False, True])
@pytest.mark.parametrize("stack", range(4))
def test_emission_order_wrappers_and_capacity(draws, relocated, stack):
    # native_routes_coarse_emit_preserves_order_wrappers_and_capacity_guards (emit.rs:5-34)
    for capacity in (0, 3, draws + 8):
        check_emit(emission_scene(draws, relocated, stack, capacity))


@pytest.mark.parametrize("mode", PAINT_MODES)
def test_paint_fast_paths(mode):
    # native_routes_coarse_emit_preserves_paint_fast_paths_and_glyph_indices (emit.rs:36-53)
    k = paint_scene(mode)
    check_emit(k)
    work, kinds = k.oracle_count()
    assert work[0] == 2


def test_nested_stack_end_order():
    # native_routes_coarse_emit_reverses_nested_stack_end_order (emit.rs:55-90)
    k = nested_end_order_scene()
    check_emit(k)
    assert k.oracle_count()[0][0] == 8


@pytest.mark.parametrize("case", range(6))
def test_mixed_stacks_keep_counts_and_wrapper_order(case):
    # native_routes_coarse_mixed_stacks_keep_matching_counts_and_wrapper_order (stack.rs:103-142)
    k = mixed_stack(case)
    work, kinds = k.oracle_count()
    assert work[0] == k.count
    assert kinds[0] == (1 if k.count == 0 else 0)
    check_emit(k)


@pytest.mark.parametrize("count", DRAWS)
@pytest.mark.parametrize("relocated", [False, True])
@pytest.mark.parametrize("stack", range(3))
@pytest.mark.parametrize("alpha", [0, 254, 255])
def test_opaque_prefix_skip(count, relocated, stack, alpha):
    # native_opaque_occlusion_preserves_allocations_pages_and_layer_semantics (occlusion.rs:81-103)
    k = opaque_scene(count, relocated, stack, alpha)
    check_emit(k)
    assert k.oracle_count()[0][0] == k.count


def test_repeated_emission_keeps_stream_origins():
    # native_opaque_occlusion_keeps_reused_preallocated_stream_origins (occlusion.rs:162-180):
    # emitting the same tile three times leaves the same stream and kind.
    k = opaque_scene(257, True, 0, 255)
    first = k.oracle_emit()
    for _ in range(2):
        again = k.oracle_emit()
        assert np.array_equal(again[0], first[0]) and again[1] == first[1]
    assert first[1] == 2 | (128 << 3)
