"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle
asan` links oracle/sanitize_driver.cpp with oracle/oracle.cpp (host code, g++), and
the driver runs every entry point on seeded scenes of each item kind against brute
force; any sanitizer report or mismatch makes it exit non-zero."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_oracle_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    r = subprocess.run([os.path.join(ORACLE, "_san", "oracle_sanitized")], capture_output=True, text=True,
                       timeout=600, env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
                                         "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 mismatches" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
