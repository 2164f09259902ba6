"""The reference's coarse count / emit known-answer scenes on the device through
tk_coarse (count, prefix and emit of every batch in one call); the scenes and
their expected records are in coarse_kats.py (src/native/runtime/tests/
coarse_count/{scene,emit,emit_scene,stack,occlusion}.rs).  The device computes
its own tile offsets, so its stream for the scene's batch starts at record 0;
its count, the records it writes (those below the capacity), the tile kind and
the overflow flag are checked."""
import numpy as np
import pytest

from coarse_kats import PAINT_MODES, count_scene, emission_scene, mixed_stack, nested_end_order_scene, opaque_scene, \
    paint_scene

pytestmark = pytest.mark.gpu

DRAWS = [0, 1, 255, 256, 257, 513]


def check_device(ctx, k, capacity=None):
    cap = k.capacity if capacity is None else capacity
    n, got, kind, status, (start, end, total) = k.device(ctx, cap)
    assert n == k.count
    assert (start, end) == (0, n)
    want = [[int(v) for v in r] for r in k.device_stream[:max(0, cap)]]
    assert got.tolist() == want
    assert kind == k.kind
    assert bool(status[0] & 1) == (total > cap)


@pytest.mark.parametrize("draws", DRAWS)
@pytest.mark.parametrize("relocated", [False, True])
@pytest.mark.parametrize("stack", range(4))
def test_count_and_emission(ctx, draws, relocated, stack):
    for capacity in (0, 3, draws + 8):
        k = emission_scene(draws, relocated, stack, capacity)
        assert k.count == count_scene(draws, relocated, stack).count
        check_device(ctx, k)


@pytest.mark.parametrize("mode", PAINT_MODES)
def test_paint_fast_paths(ctx, mode):
    check_device(ctx, paint_scene(mode))


def test_nested_stack_end_order(ctx):
    check_device(ctx, nested_end_order_scene())


@pytest.mark.parametrize("case", range(6))
def test_mixed_stacks(ctx, case):
    check_device(ctx, mixed_stack(case))


@pytest.mark.parametrize("count", DRAWS)
@pytest.mark.parametrize("relocated", [False, True])
@pytest.mark.parametrize("stack", range(3))
@pytest.mark.parametrize("alpha", [0, 254, 255])
def test_opaque_prefix_skip(ctx, count, relocated, stack, alpha):
    check_device(ctx, opaque_scene(count, relocated, stack, alpha))
