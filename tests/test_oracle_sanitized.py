"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code, gcc): every
function of oracle/oracle.c runs on exactly-sized buffers at the tier boundaries and the integer
edge values (oracle/sanitize_check.c); any out-of-bounds access, use after free or undefined
behaviour (signed overflow in the wrapping restatements, a bad shift, ...) aborts the run."""
import shutil
import subprocess

import pytest

from oracle import build as ob


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not on PATH")
def test_oracle_clean_under_asan_and_ubsan():
    exe = ob.build_sanitized(verbose=False)
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1", "PATH": "/usr/bin:/bin"}
    proc = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600, env=env)
    assert proc.returncode == 0, proc.stderr[-4000:]
    assert "runtime error" not in proc.stderr and "AddressSanitizer" not in proc.stderr, proc.stderr[-4000:]
    assert "oracle sanitizer check: ok" in proc.stdout
