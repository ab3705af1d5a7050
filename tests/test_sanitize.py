"""The CPU oracle and the library's host tree stages under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/sanitize/Makefile builds both drivers with g++ and runs
them (the host driver also requires the tree stages' labels, strengths and persistence to
equal the oracle's bit for bit)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.slow
@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_oracle_and_host_stages_under_sanitizers():
    r = subprocess.run(["make", "-C", HERE, "run"], capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "oracle stages ran clean" in r.stdout
    assert "host tree stages ran clean" in r.stdout
