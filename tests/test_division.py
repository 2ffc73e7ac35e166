"""The force pass's shared-reciprocal division equals IEEE division on the operands it is used for.

When every operand of a step is in range, the force pass computes its seven quotients per neighbour
from two correctly rounded reciprocals: a / b = RN(q0 + r y) with y = RN(1 / b), q0 = RN(a y),
r = a - b q0 (one fma).  The range is positions, velocities and pressures 0 or in [2^-40, 2^40],
densities in [2^-40, 2^40]; the density pass checks it.  Markstein's theorem says the result is the
correctly rounded quotient when nothing underflows or overflows.  ``div_identity.c`` checks this
with the CPU's IEEE division and fma:

* exhaustively, every significand of a against divisor significands near 1, near sqrt(2) and near 2,
  where the reciprocal's rounding error is largest relative to the quotient's;
* every significand of a against 4096 random divisor significands (2^35 pairs; the quotient's
  rounding depends only on the two significands, so this samples the whole divisor range);
* on random operands drawn the way the kernel forms them (sums and differences of in-range values;
  r / d with d = |r| + 1e-16).

The GPU parity tests then compare the force pass, which takes this path on every in-range step,
with the oracle's IEEE divisions bit for bit.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).with_name("div_identity.c")


@pytest.fixture(scope="module")
def div_identity(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = tmp_path_factory.mktemp("div") / "div_identity"
    base = ["gcc", "-O2", "-ffp-contract=off", "-fopenmp", str(SRC), "-o", str(exe), "-lm"]
    # hardware fma (correctly rounded either way; the library's software fma is only slower)
    if subprocess.run(base[:1] + ["-march=native"] + base[1:], capture_output=True).returncode != 0:
        subprocess.run(base, check=True)
    return exe


def _run(exe, *args):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300)
    total, bad = map(int, r.stdout.split())
    return r.returncode, total, bad


@pytest.mark.parametrize("lo,hi", [(0x0, 0x100), (0x350400, 0x350500), (0x7FFF00, 0x800000)],
                         ids=["near-1", "near-sqrt2", "near-2"])
def test_exhaustive_significands(div_identity, lo, hi):
    rc, total, bad = _run(div_identity, "exhaustive", hex(lo), hex(hi))
    assert total == (hi - lo) << 23
    assert bad == 0 and rc == 0


def test_sampled_divisor_significands(div_identity):
    rc, total, bad = _run(div_identity, "sampled", 4096, 20240611)
    assert total == 4096 << 23
    assert bad == 0 and rc == 0


def test_random_kernel_operands(div_identity):
    rc, total, bad = _run(div_identity, "random", 5_000_000, 12345)
    assert total == 20_000_000
    assert bad == 0 and rc == 0


def test_checker_detects_a_wrong_quotient(tmp_path):
    """The comparison is live: a quotient without the fma correction step differs somewhere."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = SRC.read_text().replace("return fmaf(r, y, q0);", "return q0 + 0.0f * r;")
    bad_src = tmp_path / "div_bad.c"
    bad_src.write_text(src)
    exe = tmp_path / "div_bad"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(bad_src), "-o", str(exe), "-lm"], check=True)
    rc, total, bad = _run(exe, "random", 200_000, 1)
    assert bad > 0 and rc == 1
