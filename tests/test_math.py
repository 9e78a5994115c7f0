"""Arithmetic helpers of the stage kernels (ecogen_amd/csrc/eco_math.h),
compiled for the host from the same source:

* qdiv: several IEEE quotients through one shared reciprocal,
  q = RN(a*y), r = fma(-b, q, a), RN(q + r*y), y = RN(1/b) -- must equal a/b
  bit for bit on random and adversarial operands (significands near all ones,
  near-power-of-two divisors);
* pow_pos: the isentrope's power x^gamma for x > 0 -- within 2 ulp (gamma
  1.4) and 5 ulp (gamma 4.4) of the C library's pow on x in [1e-3, 1e3], and
  the library's values for zero, negative, infinite and NaN bases."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "math_check.cpp"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        pytest.skip("no C++ compiler")
    out = tmp_path_factory.mktemp("math") / "math_check"
    subprocess.run([cxx, "-O2", "-ffp-contract=off", "-std=c++17", str(SRC), "-o", str(out)], check=True)
    return out


def _run(exe, *args):
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, check=True).stdout.split()


@pytest.mark.parametrize("seed", [1, 0xEC0C, 0x5EED5EED])
def test_qdiv_bit_identical(exe, seed):
    assert int(_run(exe, "qdiv", 5000000, seed)[0]) == 0


def test_plain_reciprocal_product_is_not_enough(exe):
    """The check is sensitive: a*RN(1/b) alone differs from a/b in a large fraction of cases."""
    assert int(_run(exe, "plain", 200000, 7)[0]) > 10000


@pytest.mark.parametrize("gamma,bound", [(1.4, 2.0), (4.4, 5.0)])
def test_pow_pos_ulp_bound(exe, gamma, bound):
    mx, mean = map(float, _run(exe, "pow", 3000000, gamma, 11))
    assert mx <= bound and mean < 0.6, (mx, mean)
