"""The independent-scene workload at full size: 512 scenes, canvas side
log-uniform in 64-2048 px, aspect 0.5-2, path counts lognormal(ln 1500, 1.2)
clipped to 10-20 000, each recorded by the clipped-paths generator.  The whole
set renders as one batch; every scene's batch image equals its own single-scene
render, and the smallest scenes equal the oracle's render bit for bit."""
import numpy as np
import pytest

import oracle
import tileink_amd as tk
from tileink_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(1500)
def test_512_scene_batch_equals_single_renders_and_oracle(ctx):
    sizes = scenes.scene_set_sizes(512)
    arrays = scenes.scene_set(sizes)
    paths = sum(a.path_count for a in arrays)
    print(f"scene set: {len(arrays)} scenes, {paths} paths, {sum(a.line_count for a in arrays)} lines", flush=True)
    assert paths == sum(n for _, _, n in sizes)
    encoded = [a.to_scene() for a in arrays]
    batch = tk.Batch(ctx, encoded)
    batch.render()
    stats = batch.frame.stats()
    assert stats["overflow"] == 0
    print(f"scene set: batch rendered, {stats['ptcl_records']} PTCL records", flush=True)
    for i, scene in enumerate(encoded):
        single = tk.Frame(ctx, scene)
        single.render()
        assert single.stats()["overflow"] == 0, i
        expected = single.image()
        del single
        got = batch.image(i)
        assert got.shape == (scene.height, scene.width)
        assert np.array_equal(got, expected), (i, sizes[i], int((got != expected).sum()))
        if i % 64 == 63:
            print(f"scene set: {i + 1} scenes equal their single renders", flush=True)
    # The oracle on the smallest scenes (by pixels x paths), whole frame.
    order = np.argsort([w * h * max(n, 1) for w, h, n in sizes])
    for i in order[:12]:
        ref, _ = oracle.render_frame(arrays[i], outputs=False)
        assert np.array_equal(batch.image(int(i)), ref), (int(i), sizes[i])
