"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only, gcc): `make -C oracle sanitize` builds oracle/lbm_oracle.c
with oracle/sanitize_check.c, which drives every oracle entry point on small
decks (every RKMETHOD, WENO5 and NND, periodic, mirrored and clamped edges,
diffuse walls, permanent cells, viscosity laws, non-uniform spacing). Any
out-of-bounds access, use after free, leak or undefined arithmetic aborts
the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-B", "-C", ORACLE, "sanitize"], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "sanitize_check")], capture_output=True, text=True, env=env,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
