"""The CPU restatement (oracle/sfb_oracle.c) and the host-only parts of the C++
API (catana_amd/csrc/catana_api.cpp) built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on small inputs (oracle/Makefile, target
sanitize): any memory error, leak (oracle) or undefined behaviour aborts the
driver, and the drivers' own consistency checks must pass."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.timeout(600)
def test_sanitized_builds_run_clean():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("no host compiler")
    if not os.path.exists(os.path.join(ROOT, "catana_amd", "libcatana_hip.so")):
        pytest.skip("libcatana_hip.so is not built")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "sanitize_oracle: all checks passed" in out
    assert "sanitize_host: all checks passed" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out
