"""The run-time mixed-radix FFT of csrc/fft_any.hpp, run on the CPU.

tests/host/fft_any_check.cpp executes the engine's stages on the host, one
emulated thread after another, for several thread counts, and compares forward
and inverse transforms with a direct FP64 DFT.  The lengths cover every
butterfly (radix 2, 3, 4, 5, 8), direct prime stages (7, 17, 67) and the sizes
of real scans: detector widths 1536, 1608, 2500 and the oversampled Radon
grids 2n = 3216, 5000, 6528.  Bar: relative L2 <= 5e-7 (FP32 arithmetic,
FP64 twiddles).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "fft_any_check.cpp")
CSRC = os.path.join(ROOT, "holotomocupy_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
LENGTHS = [2, 3, 5, 7, 8, 12, 16, 30, 60, 67, 96, 100, 134, 17, 1000, 1536, 1608, 2048, 2500, 3216, 5000, 6528]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("fftany") / "fft_any_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I" + CSRC, SRC, "-o", exe],
                   check=True, capture_output=True)
    return exe


def test_any_length_fft_matches_direct_dft(checker):
    out = subprocess.run([checker] + [str(n) for n in LENGTHS], capture_output=True, text=True, check=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    assert [int(r[0]) for r in rows] == LENGTHS
    for r in rows:
        n, fwd, inv, radices = int(r[0]), float(r[1]), float(r[2]), [int(x) for x in r[3:]]
        assert fwd < 5e-7 and inv < 5e-7, r
        prod = 1
        for x in radices:
            prod *= x
        assert prod == n, r


def test_factorisation_orders_radices(checker):
    out = subprocess.run([checker, "6528", "2500", "1536"], capture_output=True, text=True, check=True).stdout
    rad = {int(r.split()[0]): [int(x) for x in r.split()[3:]] for r in out.strip().splitlines()}
    assert rad[6528] == [8, 8, 2, 3, 17]
    assert rad[2500] == [4, 5, 5, 5, 5]
    assert rad[1536] == [8, 8, 8, 3]
