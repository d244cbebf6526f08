"""AddressSanitizer + UndefinedBehaviorSanitizer runs of host code (tests/sanitize/Makefile, run by
the build): the oracle's restatement through every entry point at both tile sizes, and the C17
host layer of the library on its GPU-free paths (kraus_to_superop, state_len, rejection of malformed
state files).  A sanitizer report makes the program exit non-zero and print its diagnosis."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def run(name, *args):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (make -C tests/sanitize)")
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([path, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    return r.stdout


@pytest.mark.parametrize("tile", ["t3", "t5"])
def test_oracle_under_sanitizers(tile):
    assert "oracle calls" in run(f"san_oracle_{tile}")


def test_host_layer_under_sanitizers(tmp_path):
    assert "san_host ok" in run("san_host", str(tmp_path))
