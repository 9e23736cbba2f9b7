"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): every entry
point (quadrature, each model on its reference grid, brute force, cutoff, periodic,
exclusions, the general-expression integral chain, grid inspection) on small deterministic
systems, built by `make -C oracle sanitize`; any memory or undefined-behaviour finding aborts
the program, and its dense (reference loop order) and sparse forms must agree."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_clean_under_asan_and_ubsan():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc/make not available")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "_asan", "sanitize_check")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all oracle entry points clean" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
