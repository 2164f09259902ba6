"""bench.py's host-side contract: the one-GPU guard, the algorithmic byte count
of a fine launch, and the measured-traffic record keyed to the kernel source."""
import os
import subprocess
import sys
import types

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_fine_algorithmic_bytes_closed_form():
    # 24 B per walked record, 20 B per used segment, 12 B per batch x tile entry,
    # 4 B per pixel written
    scene = types.SimpleNamespace(batch_count=3, tiles_width=4, tiles_height=2, width=64, height=32)
    assert bench.fine_algorithmic_bytes(10, 7, scene) == 24 * 10 + 20 * 7 + 12 * 3 * 4 * 2 + 4 * 64 * 32


def test_measured_traffic_only_for_the_profiled_source(monkeypatch):
    got = bench.measured_traffic("tk_k_fine")
    assert got is None or got > 0
    monkeypatch.setattr(bench, "source_digest", lambda: "not-the-profiled-source")
    assert bench.measured_traffic("tk_k_fine") is None
    assert bench.measured_traffic("no_such_kernel") is None
