"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
oracle/sanitize_check.c calls every oracle entry (kernels of all orders,
the eight varifold kernel pairs with and without gradient, objgrd with the
varifold and l2 matching terms, the optimiser pieces and a registration) on
small 2D and 3D problems; the build (oracle/Makefile `sanitize`) aborts on
the first report."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("gcc") is None, reason="needs make and gcc")
def test_oracle_is_clean_under_asan_and_ubsan():
    b = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    # verify_asan_link_order=0: a library the environment preloads may come
    # before the ASan runtime in the link order
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "sanitize_check")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "sanitize check done" in r.stdout
