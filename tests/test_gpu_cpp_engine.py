"""Runs the C++ engine-mirror test program (tests/cpp/engine_test.cpp): the reference's StatevectorEngineTest,
StatevectorNoiseTest and StatevectorAlgorithmSuiteTest cases against HipStatevectorEngine."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent.parent / "statecraft_amd" / "lib" / "engine_test"


def test_engine_binary_links():
    assert BIN.exists(), "build with python -m statecraft_amd.build"
    out = subprocess.run(["ldd", str(BIN)], capture_output=True, text=True).stdout
    assert "libstatecraft_engine.so" in out and "not found" not in out


@pytest.mark.gpu
def test_cpp_engine_suite():
    res = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failure(s)" in res.stdout
