"""The device kernels form several IEEE quotients with one shared reciprocal
(qdiv in ecogen_amd/csrc/eco_physics.hpp): q = RN(a*y), r = fma(-b, q, a),
RN(q + r*y), y = RN(1/b).  This pins that construction to the correctly
rounded quotient a/b, bit for bit, on random and adversarial operands
(significands near all ones, near-power-of-two divisors), compiled with the
host C compiler and its fused multiply-add."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "qdiv_check.c"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    out = tmp_path_factory.mktemp("qdiv") / "qdiv_check"
    subprocess.run([cc, "-O2", "-ffp-contract=off", str(SRC), "-o", str(out), "-lm"], check=True)
    return out


@pytest.mark.parametrize("seed", [1, 0xEC0C, 0x5EED5EED])
def test_qdiv_bit_identical(exe, seed):
    r = subprocess.run([str(exe), "5000000", str(seed)], capture_output=True, text=True, check=True)
    assert int(r.stdout.strip()) == 0


def test_plain_reciprocal_product_is_not_enough(exe, tmp_path):
    """The check is sensitive: a*RN(1/b) alone differs from a/b in a large
    fraction of cases (the correction step is what makes it exact)."""
    src = SRC.read_text().replace("return fma(r, y, q);", "(void)r; return q;")
    p = tmp_path / "plain.c"
    p.write_text(src)
    cc = shutil.which("gcc") or shutil.which("cc")
    out = tmp_path / "plain"
    subprocess.run([cc, "-O2", "-ffp-contract=off", str(p), "-o", str(out), "-lm"], check=True)
    r = subprocess.run([str(out), "200000", "7"], capture_output=True, text=True, check=True)
    assert int(r.stdout.strip()) > 10000
