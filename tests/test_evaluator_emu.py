"""The product's bit-plane evaluator, compiled for the host and run through a whole-grid CPU
window (tests/cpp/evaluator_emu.cpp), equals the oracle's standard evaluator for every rule,
both layouts and both step parities. Catches evaluator bugs without a GPU."""
import subprocess
from pathlib import Path

REPO = Path(__file__).resolve().parent.parent


def test_evaluator_on_cpu(tmp_path):
    exe = tmp_path / "evaluator_emu"
    subprocess.run(["g++", "-std=c++20", "-O1", f"-I{REPO}/include", f"-I{REPO}/oracle",
                    str(REPO / "tests/cpp/evaluator_emu.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout
