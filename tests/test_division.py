"""The kernels divide by values shared across many quotients (the init's
dt + 1, a channel's sigma, the S/N divisor beyond the mask extent) with a
rounded reciprocal and one FMA correction instead of a full division.  This
checks, on the CPU and bit for bit, that the sequence reproduces IEEE x / y
for every x mantissa, every odd integer divisor up to 1023, 64 random float
divisors and sqrtf(2) (the boxcar's S/N divisor for odd widths, up to a power
of two; the exhaustive run over all odd divisors below 32768 gives no
mismatch either)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_shared_reciprocal_division_is_exact(tmp_path):
    exe = str(tmp_path / "div_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", os.path.join(HERE, "cpp", "div_check.c"), "-o", exe,
                    "-lm"], check=True)
    out = subprocess.run([exe, "1023", "64"], check=True, stdout=subprocess.PIPE, text=True).stdout
    assert int(out.strip()) == 0
