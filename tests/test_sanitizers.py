"""CPU sanitizer builds (no GPU needed).

- The oracle's own tests run against an AddressSanitizer +
  UndefinedBehaviorSanitizer build of the oracle (oracle/_build/liborc_san.so),
  the sanitizer runtimes preloaded into the Python process; a bad input planted
  in one oracle call shows that the sanitizers are live in that build.
- The host code of libpbf.so (parameter validation, the grid and kernel
  constants derived inside pbf_create before its first HIP call, the status
  strings, the null-context refusals) runs under the same sanitizers:
  tests/cpp/pbf_host_check.cpp against a build of the library whose host code
  is instrumented (physika-old_amd/_build_san/).
Any sanitizer report aborts the process under test."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SAN = ROOT / "oracle" / "_build" / "liborc_san.so"
HOST_CHECK = ROOT / "physika-old_amd" / "_build_san" / "pbf_host_check"
ORACLE_TESTS = ["test_oracle_grid_neighbors.py", "test_oracle_physics.py", "test_oracle_adjoint.py",
                "test_oracle_boundary_viscosity.py", "test_oracle_mass_fixed.py", "test_golden.py"]
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def _gcc_runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def _oracle_env():
    asan, ubsan = _gcc_runtime("libasan.so"), _gcc_runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not found")
    env = dict(os.environ)
    env.update(LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", PBF_ORACLE_LIB=str(ORACLE_SAN),
               OMP_NUM_THREADS="4")
    return env


@pytest.fixture(scope="module")
def oracle_san():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "sanitize"], check=True, capture_output=True)
    return _oracle_env()


def test_oracle_tests_under_sanitizers(oracle_san):
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-x"]
                       + [str(ROOT / "tests" / t) for t in ORACLE_TESTS],
                       env=oracle_san, capture_output=True, text=True, timeout=900, cwd=str(ROOT))
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert not any(rep in out for rep in REPORTS), out[-4000:]
    assert " passed" in out


def test_oracle_sanitizer_build_is_live(oracle_san):
    """A CSR whose last offset runs past the elements array makes the oracle's
    density read out of bounds: the instrumented build must catch it."""
    code = f"""
import sys, numpy as np
sys.path.insert(0, {str(ROOT / 'oracle')!r})
import oracle
assert str(oracle._LIB_PATH).endswith('liborc_san.so')
assert any('libasan' in line for line in open('/proc/self/maps'))
x = np.zeros((2, 3), np.float32); x[1, 0] = 0.001
off = np.array([0, 2, 40], np.int32)
el = np.array([0, 1, 0, 1], np.int32)
oracle.density(oracle.make_params(), x, off, el)
print('NOT CAUGHT')
"""
    r = subprocess.run([sys.executable, "-c", code], env=oracle_san, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "NOT CAUGHT" not in r.stdout
    assert "heap-buffer-overflow" in r.stderr, r.stderr[-3000:]


def test_library_host_code_under_sanitizers():
    if not Path("/opt/rocm/bin/hipcc").exists():
        pytest.skip("hipcc not available")
    subprocess.run(["make", "-C", str(ROOT / "physika-old_amd" / "csrc"), "-f", "sanitize.mk"], check=True,
                   capture_output=True)
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(HOST_CHECK.parent / "libpbf_san.so")],
                        capture_output=True, text=True).stdout
    assert "__asan_report" in nm and "__ubsan_handle" in nm  # the host code is instrumented
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(HOST_CHECK)], env=env, capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ALL PASSED" in r.stdout and "FAIL " not in r.stdout
    assert not any(rep in out for rep in REPORTS), out[-4000:]
