"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the oracle
and the library's host geometry builders (csrc/sphere.cpp), driven by
oracle/sanitize_check.cpp on small systems (make -C oracle sanitize).  CPU
only; any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_is_sanitizer_clean():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "SANITIZE-OK" in r.stdout
