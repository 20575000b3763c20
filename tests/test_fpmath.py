"""The library's float exp/log/pow (vdb_fpmath.h, the same header the HIP kernels include)
built for the host and compared with correctly rounded values from x87 extended precision
(expl/logl/powl, one rounding to float). A strided subset of all float inputs runs here; a
stride of 1 covers every float in the exp domain [-110, 90] and every positive finite float
for log."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nanovdb-editor_amd", "csrc")


@pytest.fixture(scope="module")
def fpmath_check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("fpmath") / "fpmath_check"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fopenmp", "-ffp-contract=off", "-I", CSRC,
                    os.path.join(ROOT, "tests", "native", "fpmath_check.c"), "-o", str(exe), "-lm"], check=True)
    return str(exe)


@pytest.mark.parametrize("fn,stride", [("exp", 61), ("log", 67)])
def test_correctly_rounded(fpmath_check, fn, stride):
    out = subprocess.run([fpmath_check, fn, str(stride)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    checked = int(out.stdout.split()[2])
    assert checked > 1_000_000


def test_pow_correctly_rounded(fpmath_check):
    out = subprocess.run([fpmath_check, "pow", "1", "300000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
