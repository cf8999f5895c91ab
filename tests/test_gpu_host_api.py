"""Runs the C++ host-API test program (tests/cpp/host_api_test.cpp) on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
BIN = Path(__file__).parent / "cpp" / "host_api_test"


def test_cpp_host_api():
    if not BIN.exists():
        subprocess.run(["make", "-C", str(BIN.parent)], check=True)
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ALL OK" in r.stdout
