"""bench.py's contract on the CPU: one GPU only, the algorithmic byte count of
the step, the committed traffic summary it reads, and the rate-probe library
it loads (symbols only; no GPU call)."""
import ctypes as C
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_of_the_step():
    # per particle: sort 60, integrate 48, K x (32 + 40), velocity 48; 12 per
    # grid cell; no neighbor-list bytes (neighbor data count as on-chip)
    n, c, k = 1000, 77, 5
    assert bench.algorithmic_bytes(n, c, k) == (60 + 48 + 72 * k + 48) * n + 12 * c
    assert bench.algorithmic_bytes(n, c, 5) == 516 * n + 12 * c
    # the 1M lattice on the 85^3 reference grid: 548.4 MB
    assert abs(bench.algorithmic_bytes(1048576, 85 ** 3, 5) - 548.4e6) < 0.1e6


def test_algorithmic_flops_of_the_step():
    # 9 per pair (distance test) + 57 per pair and iteration + 30 per particle
    n, p = 1000, 70000
    assert bench.algorithmic_flops(n, p, 5) == (9 + 57 * 5) * p + 30 * n
    assert bench.algorithmic_flops(n, p, 0) == 9 * p + 30 * n


def test_kernel_counts_and_bound():
    n, p = 1 << 20, 71_000_000
    assert bench.kernel_counts("k_density_lambda", n, p) == (32 * n, 33 * p)
    assert bench.kernel_counts("k_displacement", n, p) == (40 * n, 24 * p)
    assert bench.kernel_counts("k_neighbors", n, p) == (16 * n, 9 * p)
    # 41.9 MB and 1.70 GFLOP in 72.8 us: VALU-bound at 0.149 of 157.3 TFLOP/s
    b, f = bench.kernel_counts("k_displacement", n, p)
    r = bench.roofline_of(b, f, 72.8e-6)
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s" and r["peak"] == 157.3
    assert abs(r["frac"] - f / 72.8e-6 / 157.3e12) < 1e-4 and 0.14 < r["frac"] < 0.16
    assert abs(r["hbm_frac"] - b / 72.8e-6 / 8e12) < 1e-4
    # a copy-like kernel is HBM-bound
    r = bench.roofline_of(1e9, 1e6, 1e9 / 4e12)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and abs(r["frac"] - 0.5) < 1e-4


def test_baseline_threads_follow_the_environment(monkeypatch):
    aff = len(__import__("os").sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.baseline_threads() == 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.baseline_threads() == aff
    info = bench.host_cpu()
    assert info["affinity_cpus"] == aff and isinstance(info["model"], str)


def test_committed_traffic_summary_is_read():
    for k in ("k_density_lambda", "k_displacement", "k_neighbors"):
        t, name = bench.profiled_traffic(k)
        assert isinstance(t, int) and 1e7 < t < 1e9, (k, t)
        assert name == k or name.startswith(k + "<") or name.startswith(k + "_p<")
    assert bench.profiled_traffic("no_such_kernel") == (None, None)


def test_rate_probe_library_exports():
    import __graft_entry__ as G
    so = Path(G.PKG_DIR) / "_build" / "libpbf_peaks.so"
    if not so.exists():
        pytest.skip("libpbf_peaks.so not built")
    lib = C.CDLL(str(so))
    assert hasattr(lib, "pbf_peak_copy") and hasattr(lib, "pbf_peak_fma")
