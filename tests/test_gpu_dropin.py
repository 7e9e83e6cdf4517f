"""GPU: the C++ drop-in classes (FMM / NaiveUpdate / SyplecticEuler /
RungeKutta4 behind the reference's SystemUpdateMethod API) running the
reference's own test scenarios (tests/cpp/test_dropin.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nbody-fmm_amd", "_build", "test_dropin")

pytestmark = pytest.mark.gpu


def test_dropin_program(tmp_path):
    assert os.path.exists(EXE), "build first: python nbody-fmm_amd/build.py"
    r = subprocess.run([EXE, str(tmp_path / "frames")], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
