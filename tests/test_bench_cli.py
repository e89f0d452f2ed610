"""bench.py's command line on the CPU: the one-GPU contract and the options
the committed profiles use (no GPU work is started)."""

import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_help_lists_the_profiling_options():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--precision", "--ensemble-boxes", "--accuracy-steps", "--cpu-baseline", "--ensemble-tiling"):
        assert opt in r.stdout


def test_ensemble_tiling_takes_four_values():
    r = run("--gpus", "2", "--ensemble-tiling", "32", "8", "16")
    assert r.returncode == 2 and "expected 4 arguments" in r.stderr


def test_cpu_baseline_ranks_follow_the_host(monkeypatch):
    """The reference baseline uses the largest rank count built that fits the
    physical cores in the affinity mask and the CPU share (OMP_NUM_THREADS),
    and reports the CPU model."""
    sys.path.insert(0, REPO)
    import bench
    monkeypatch.setenv("OMP_NUM_THREADS", "16")
    h = bench.host_cores()
    assert 1 <= h["cores"] <= 16 and h["cores"] <= h["physical_cores_in_affinity"]
    assert h["model"] and h["cpu_share"] == 16
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_cores()["cpu_share"] is None
    exe, ranks = bench.baseline_exe(16)
    if exe is not None:  # built where the reference sources exist
        assert ranks == 16 and exe.endswith("bench512_r16/3dmhd.exe")
        assert bench.baseline_exe(12)[1] == 8 and bench.baseline_exe(4) == (None, 0)
