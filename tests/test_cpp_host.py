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
    for prog in ("test_gpu_utils", "test_corrections", "test_imager", "test_fits",
                 "test_div_rn"):
        assert os.path.exists(os.path.join(BIN, prog)), prog


def test_cpp_fits_io():
    """FITS reader / writer and the correlation-matrix conversion (no GPU)."""
    exe = os.path.join(BIN, "test_fits")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "All tests passed." in r.stdout


def test_cpp_div_rn():
    """The gridding plans' reciprocal-based FP64 division equals IEEE division
    on 2e7 random operand pairs (no GPU)."""
    r = subprocess.run([os.path.join(BIN, "test_div_rn")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "All tests passed." in r.stdout


def test_fits_reader_matches_numpy():
    """The C++ reader against an independent parse of the same file: 2880-byte
    header blocks up to END, then big-endian float64 samples."""
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "eda2_chan204",
                        "chan_204_20230205T034618_simul_vis_XX_CorrMatrix_RE.fits")
    raw = open(path, "rb").read()
    off = 0
    while True:
        block = raw[off:off + 2880]
        off += 2880
        cards = [block[i:i + 80].decode("ascii") for i in range(0, 2880, 80)]
        if any(c.startswith("END ") or c.rstrip() == "END" for c in cards):
            break
    data = np.frombuffer(raw, dtype=">f8", count=256 * 256, offset=off)
    r = subprocess.run([os.path.join(BIN, "test_fits"), path], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    nx, ny, total, first, last = r.stdout.split()
    assert (int(nx), int(ny)) == (256, 256)
    assert float(first) == data[0] and float(last) == data[-1]
    assert abs(float(total) - float(np.sum(data))) <= 1e-9 * float(np.sum(np.abs(data)))
