"""bench.py's command line and workload definitions (no GPU): one GPU only, the workload
sizes of the benchmark configurations and their algorithmic flop counts."""

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          env=e, timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2 and "one GPU" in r.stderr
    r = run_bench(env={"WORLD_SIZE": "4"})
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_the_workloads():
    r = run_bench("--help")
    assert r.returncode == 0
    for w in ("dba", "fodo", "da", "diff", "batched"):
        assert w in r.stdout
    assert "--streamed-turns" in r.stdout


def test_workload_sizes_and_flops():
    import bench
    from trackpad_hip.flops import flops_per_turn

    beam, lat, n, T, mk, ap = bench.workload("dba", 0, 0)
    assert (n, T, len(lat), ap) == (1_000_000, 1000, 501, None) and flops_per_turn(lat) == 313_587
    beam, lat, n, T, mk, ap = bench.workload("fodo", 0, 0)
    assert (n, T, len(lat), ap) == (10_000, 1000, 32, 0.02) and flops_per_turn(lat) == 7_472
    beam, lat, n, T, mk, ap = bench.workload("da", 0, 0)
    assert (n, T, len(lat)) == (3 * 1024 * 1024, 20, 2001) and flops_per_turn(lat) == 3_299_127
    s = mk(2048)
    # a spread sample of the grid: every momentum, the full square
    assert s.shape == (2048, 6) and set(np.unique(s[:, 5])) == {-0.02, 0.0, 0.02}
    assert s[:, 0].min() == pytest.approx(-0.015) and s[:, 2].max() == pytest.approx(0.015, abs=1e-4)
    beam, lat, n, T, mk, ap = bench.workload("da_lossless", 0, 0)
    assert (n, T, len(lat)) == (1_000_000, 5, 2001) and mk(8).shape == (8, 6)
    beam, lat, n, T, mk, ap = bench.workload("dba", 4096, 100)
    assert (n, T) == (4096, 100) and mk(16).shape == (16, 6)


def test_help_lists_the_metric_legs():
    r = run_bench("--help")
    assert r.returncode == 0
    for opt in ("--parity-turns", "--compare-variants", "--no-batched", "--batched-ckpt-stride"):
        assert opt in r.stdout


def test_cpu_baseline_threads(monkeypatch):
    """The CPU baseline runs on every core the process may use, or on the share that
    OMP_NUM_THREADS grants it, and says which."""
    import bench

    vis = bench.visible_cores()
    monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
    assert bench.host_cores() == vis and "all visible" in bench.cores_note(vis)
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.host_cores() == 1
    if vis > 1:
        assert "OMP_NUM_THREADS" in bench.cores_note(1)
    assert bench.BOUND == "vector_fp64"
