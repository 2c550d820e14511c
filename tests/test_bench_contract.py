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
    # per particle: sort 60, integrate 48, K x (32 + 40), velocity 48; 12 per cell;
    # the 2-byte list is written once and read by 2K passes
    n, c, p, k = 1000, 77, 5000, 5
    assert bench.algorithmic_bytes(n, c, p, k) == (60 + 48 + 72 * k + 48) * n + 12 * c + 2 * p * (2 * k + 1)
    assert bench.algorithmic_bytes(n, c, p, 5) == 516 * n + 12 * c + 22 * p


def test_committed_traffic_summary_is_read():
    for k in ("k_density_lambda", "k_displacement"):
        t = bench.profiled_traffic(k)
        assert isinstance(t, int) and 1e8 < t < 1e9, (k, t)
    assert bench.profiled_traffic("no_such_kernel") is None


def test_rate_probe_library_exports():
    import __graft_entry__ as G
    so = Path(G.PKG_DIR) / "_build" / "libpbf_peaks.so"
    if not so.exists():
        pytest.skip("libpbf_peaks.so not built")
    lib = C.CDLL(str(so))
    assert hasattr(lib, "pbf_peak_copy") and hasattr(lib, "pbf_peak_fma")
