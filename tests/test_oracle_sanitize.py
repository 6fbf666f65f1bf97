"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle sanitize`
builds rd_oracle.cpp with -fsanitize=address,undefined (aborting on the first report) into a
program that drives every entry point on seeded 2-D and 3-D clouds in FP32 and FP64."""

import os
import shutil
import subprocess

import pytest

from conftest import ROOT

ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_oracle_is_clean_under_sanitizers():
    r = subprocess.run(["make", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "oracle_selftest")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "oracle self-test ok" in r.stdout
