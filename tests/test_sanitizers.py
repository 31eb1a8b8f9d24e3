"""Host sanitizer runs (CPU): the CPU oracle and the C-ABI's host-side descriptor handling and
argument validation built with AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code
only; tests/asan/Makefile) and driven over edge cases (delays outside the sample range,
sparse indices outside [0, T), ranks 1-6, zero, huge and negative dims and strides, every
type, compute type and algorithm).  A sanitizer report fails the run."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "asan"


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    r = subprocess.run(["make", "-C", str(HERE), "-j4"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return HERE / "_build"


@pytest.mark.parametrize("exe, marker", [("oracle_asan", "cases clean"),
                                         ("plan_asan", "validations clean")])
def test_sanitized_host_code(built, exe, marker):
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(built / exe)], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert marker in r.stdout, out
    assert "runtime error" not in out and "AddressSanitizer" not in out, out
