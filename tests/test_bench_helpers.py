"""The benchmark script's host-side accounting, on the CPU: per-shot latency
from the engine's batch events when the forward and adjoint calls run
different batch counts; the HBM-traffic figure taken only from a counter
summary of the kernel source in the tree; more than one GPU refused."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_shot_latency_with_unequal_batch_counts():
    # forward: 10 batches of 8 shots (1 ms each); adjoint: 4 batches of 20 shots (10 ms each)
    bms = [1.0] * 10 + [10.0] * 4
    bsh = [8] * 10 + [20] * 4
    lat = bench.shot_latencies(bms, bsh, 80, 2)
    assert len(lat) == 80 and np.allclose(lat, 11.0)
    # two steps: twice the shots, each with its own batches
    bms2 = bms + [2.0] * 10 + [20.0] * 4
    lat2 = bench.shot_latencies(bms2, bsh + bsh, 80, 2)
    assert len(lat2) == 160 and np.allclose(lat2[:80], 11.0) and np.allclose(lat2[80:], 22.0)


def test_shot_latency_refuses_partial_calls():
    with pytest.raises(RuntimeError):
        bench.shot_latencies([1.0] * 9, [8] * 9, 80, 2)


def test_traffic_is_read_only_for_the_kernel_source_it_was_taken_on(tmp_path, monkeypatch):
    src = tmp_path / "kernels.hip"
    src.write_text("kernel source\n")
    sha = hashlib.sha256(src.read_bytes()).hexdigest()[:16]
    summary = tmp_path / "traffic.md"
    summary.write_text(f"# traffic\n\nKERNELS_SHA16={sha}\nTRAFFIC_BYTES_PER_LAUNCH=123456\n")
    monkeypatch.setattr(bench, "KERNEL_SRC", str(src))
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(summary))
    value, where = bench.measured_traffic()
    assert value == 123456.0 and where.endswith("traffic.md")
    src.write_text("changed kernel source\n")
    value, why = bench.measured_traffic()
    assert value is None and "another kernels.hip" in why


def test_committed_traffic_summary_names_a_kernel_source():
    text = open(bench.TRAFFIC_FILE).read()
    assert "KERNELS_SHA16=" in text and "TRAFFIC_BYTES_PER_LAUNCH=" in text


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
