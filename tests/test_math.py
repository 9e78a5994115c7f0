"""Arithmetic helpers of the stage kernels (ecogen_amd/csrc/eco_math.h),
compiled for the host from the same source:

* qdiv: several IEEE quotients through one shared reciprocal,
  q = RN(a*y), r = fma(-b, q, a), RN(q + r*y), y = RN(1/b) -- must equal a/b
  bit for bit on random and adversarial operands (significands near all ones,
  near-power-of-two divisors);
* pow_pos: the isentrope's power x^gamma for x > 0, the C library's pow
  algorithm restated -- equal to the library's pow bit for bit on x in
  [1e-3, 1e3] (weak waves near 1 oversampled) for the exponents the EOS use
  (gamma and 1/gamma of air, water, helium, oil), and the library's values for
  zero, negative, infinite and NaN bases.  The library is glibc's pow on an
  x86-64 host with FMA, the build the reference's results come from."""
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


@pytest.mark.parametrize("gamma", [1.4, 1 / 1.4, 4.4, 1 / 4.4, 1.67, 1 / 1.67, 2.35, 1 / 2.35])
def test_pow_pos_equals_library_pow(exe, gamma):
    assert int(_run(exe, "pow", 4000000, repr(gamma), 11)[0]) == 0


def test_exp_lib_equals_library_exp(exe):
    """THINC's exponential (exp_lib, the exp part of pow_pos) equals the C
    library's exp bit for bit."""
    assert int(_run(exe, "exp", 4000000, 5)[0]) == 0


def test_log_lib_is_the_correctly_rounded_log(exe):
    """The saturation laws' logarithm (log_lib: the double-double log of
    pow_pos rounded once) against the C library's log on uniformly drawn
    temperatures in [200, 2000] K and pressures in [1e3, 1e10] Pa: the two
    disagree in about one call in 10^4 to 10^5 (the library's log is within
    0.52 ulp, not correctly rounded), and wherever they disagree log_lib is
    the correctly rounded value (checked against a 200-bit logarithm)."""
    import mpmath as mp
    r = subprocess.run([str(exe), "log", "2000000", "9"], capture_output=True, text=True, check=True)
    bad = int(r.stdout.split()[0])
    assert 0 < bad < 2000000 // 5000
    mp.mp.prec = 200
    rows = [line.split() for line in r.stderr.splitlines() if line.strip()]
    assert rows
    for xs, libs, mine in rows:
        x, lib_v, mine_v = (float.fromhex(v) for v in (xs, libs, mine))
        exact = float(mp.log(mp.mpf(x)))
        assert mine_v == exact and lib_v != exact


def test_thinc_constants_are_the_library_values():
    """cosh(1.6) and tanh(1.6) of THINC (beta = 1.6, CellInterfaceO2.cpp:228-247)
    are written into the kernel as the C library's values."""
    import math
    import re
    src = (SRC.parent.parent / "ecogen_amd" / "csrc" / "eco_stage.inl").read_text()
    m = re.search(r"COSH_BETA = ([0-9a-fx.p+-]+), TANH_BETA = ([0-9a-fx.p+-]+);", src)
    assert m and "const double beta = 1.6" in src
    assert float.fromhex(m.group(1)) == math.cosh(1.6)
    assert float.fromhex(m.group(2)) == math.tanh(1.6)


def test_pow_pos_check_is_sensitive(tmp_path):
    """A power computed without the library's final fma contraction differs
    from the library in about one call in a thousand: the check above sees it."""
    src = SRC.read_text().replace("return std::fma(scale, tp, scale);", "return scale + scale * tp;")
    hdr = (SRC.parent.parent / "ecogen_amd" / "csrc" / "eco_math.h").read_text()
    hdr = hdr.replace("return std::fma(scale, tp, scale);", "return scale + scale * tp;")
    (tmp_path / "eco_math.h").write_text(hdr)
    (tmp_path / "eco_pow_tables.h").write_text(
        (SRC.parent.parent / "ecogen_amd" / "csrc" / "eco_pow_tables.h").read_text())
    (tmp_path / "m.cpp").write_text(src.replace('#include "../ecogen_amd/csrc/eco_math.h"', '#include "eco_math.h"'))
    cxx = shutil.which("g++") or shutil.which("c++")
    subprocess.run([cxx, "-O2", "-ffp-contract=off", "-std=c++17", str(tmp_path / "m.cpp"), "-o",
                    str(tmp_path / "m")], check=True)
    assert int(_run(tmp_path / "m", "pow", 400000, "4.4", 11)[0]) > 20
