"""AddressSanitizer and UndefinedBehaviorSanitizer builds, CPU only: the oracle's C under a driver that runs every
function on small and ragged inputs (oracle/Makefile, target `sanitize`), and the host code of libcubo_hip under a
driver of its host-only paths: the .npy parser over malformed files, the loader's file checks, the probe size,
availability, the phase calls' refusals (tests/sanitize/Makefile). A sanitizer report makes the driver exit
non-zero."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, timeout):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"{' '.join(args)} failed:\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
    return r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_oracle_under_sanitizers():
    out = _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], 300)
    assert "oracle sanitize driver: done" in out


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="no hipcc")
def test_library_host_code_under_sanitizers():
    out = _run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "tests", "sanitize"), "run"], 900)
    assert "host sanitize driver: done (0 unexpected)" in out
