"""The oracle under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer: `make -C oracle sanitize` builds oracle/sanitize_check.cpp,
which calls every entry point of oracle.h on small inputs (empty, one
particle, coincident particles, particles outside a fixed root, worker counts
1-5, row subsets), into oracle/_san/.  Both builds must run clean and report
every consistency check passed.  Host code only."""
import os
import subprocess

import pytest

from helpers import ROOT

ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def sanitized():
    r = subprocess.run(["make", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("build", ["check_asan", "check_tsan"])
def test_oracle_clean_under_sanitizer(sanitized, build):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([os.path.join(sanitized, build)], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "all entry points clean" in r.stdout
    assert "runtime error" not in out and "Sanitizer" not in out, out[-4000:]
