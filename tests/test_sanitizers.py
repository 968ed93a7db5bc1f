"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU: the oracle (every entry point on
small scenes of every kind) and the library's host code (argument validation, error contract,
index / state-count helpers, plan workspace layout; csrc/api.hip built with sanitized host code).
Built by tests/sanitize/Makefile into _build/san; the sanitizers abort on the first finding."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = os.path.join(HERE, "sanitize")
OUT = os.path.join(ROOT, "_build", "san")
ENV = {**os.environ, "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1"}


def _make(target):
    if not os.path.isdir(SAN) or shutil.which("make") is None or shutil.which("gcc") is None:
        pytest.skip("sanitizer build files or tools not present")
    r = subprocess.run(["make", "-s", "-j4", target], cwd=SAN, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-2000:]


def test_oracle_under_sanitizers():
    _make(os.path.join(OUT, "oracle_san"))
    r = subprocess.run([os.path.join(OUT, "oracle_san")], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0 and "oracle driver ok" in r.stdout, (r.stdout[-1000:], r.stderr[-3000:])


def test_library_host_code_under_sanitizers():
    if not os.path.exists(os.path.join(ROOT, "_build", "mrcc", "kernels.hip.o")):
        pytest.skip("the library's objects are not built (run __graft_entry__.build())")
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not present")
    _make(os.path.join(OUT, "host_san"))
    r = subprocess.run([os.path.join(OUT, "host_san")], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0 and "host driver ok" in r.stdout, (r.stdout[-1000:], r.stderr[-3000:])
