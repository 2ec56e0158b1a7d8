"""Host AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU code:
the oracle, the serving scheduler and the CPU build of the NTT tiles, built
with gcc/g++ -fsanitize=address,undefined (tests/emu/Makefile,
sanitize_check) and driven through representative calls; any sanitizer
finding aborts the run."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")


def test_cpu_code_is_clean_under_asan_and_ubsan():
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    b = subprocess.run(["make", "-s", "-C", HERE, "sanitize_check"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(HERE, "sanitize_check")], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "sanitized checks passed" in r.stdout
