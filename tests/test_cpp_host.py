"""Runs the C++ test programs (tests/cpp/*.cpp, built by imager_amd/build.py)
that exercise the C++ drop-in class CPacerImagerMI355X and the C ABI from
C++, in the shape of the reference's own test executables."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


@pytest.mark.gpu
@pytest.mark.parametrize("prog", ["test_gpu_utils", "test_corrections", "test_imager"])
def test_cpp_program(prog):
    exe = os.path.join(BIN, prog)
    assert os.path.exists(exe), f"{exe} is not built (python -m imager_amd.build)"
    print("running", prog, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "All tests passed." in r.stdout


def test_cpp_programs_are_built():
    for prog in ("test_gpu_utils", "test_corrections", "test_imager"):
        assert os.path.exists(os.path.join(BIN, prog)), prog
