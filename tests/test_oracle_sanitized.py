"""The oracle's C restatement under AddressSanitizer and UndefinedBehavior-
Sanitizer: the oracle tests (tests/test_oracle*.py) run in a child Python
process with oracle/libffo_asan.so (`make -C oracle asan`) in place of
libffo.so and libasan preloaded; any sanitizer report aborts the child, so
the bar is a clean exit with every oracle test passing.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _libasan():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                           check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(_libasan() is None, reason="gcc's libasan is not available")
def test_oracle_tests_pass_under_asan_and_ubsan():
    lib = os.path.join(ORACLE, "libffo_asan.so")
    subprocess.run(["make", "-s", "-C", ORACLE, "asan"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=_libasan(), FFO_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    tests = sorted(os.path.join(ROOT, "tests", f) for f in os.listdir(os.path.join(ROOT, "tests"))
                   if f.startswith("test_oracle") and f.endswith(".py") and f != os.path.basename(__file__))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert " passed" in r.stdout and "failed" not in r.stdout
