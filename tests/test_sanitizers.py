"""The CPU oracle and the are:: file formats (PFM dumps, accumulation
checkpoints) under AddressSanitizer and UndefinedBehaviorSanitizer: g++
builds oracle/san/san_main.cpp with them, which calls every oracle entry
point on scenes of every primitive, material and texture kind (pairwise and
greedy four-wide trees), the error paths, and feeds truncated and damaged
files to the loaders.  Any report aborts the run (-fno-sanitize-recover)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "oracle", "san")


@pytest.mark.skipif(not os.path.isdir(SAN) or shutil.which("g++") is None, reason="needs oracle/san and g++")
def test_oracle_and_formats_are_clean_under_asan_and_ubsan(tmp_path):
    b = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "build", "orc_san"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "san: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
