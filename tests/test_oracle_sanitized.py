"""The C++ restatements of the oracle (tile, kNN, import, peaks) under
AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle sanitize`
builds them with -fsanitize=address,undefined -fno-sanitize-recover=all, and
the oracle's own CPU tests run against that build in a child Python with
libasan preloaded, so an out-of-bounds access, use after free or undefined
operation ends the run with an error."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = ["tests/test_tile_oracle.py", "tests/test_knn_oracle.py", "tests/test_import_oracle.py",
         "tests/test_peaks_oracle.py", "tests/test_genes_oracle.py"]


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_oracle_cpp_under_asan_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("no libasan for g++ on this host")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env.update(SNAPGPU_ORACLE_SANITIZED="1", LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider", *TESTS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in tail and "runtime error" not in tail, tail
