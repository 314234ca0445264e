"""The ordered kernels divide by a distance and by a mass through the correctly rounded
reciprocal y = 1/b and two residual corrections q <- q + (a - b q) y (each an FMA):
bitwise parity with the reference's divisions needs that to be the correctly rounded
quotient.  This restates the sequence in C on the host (gcc, the C library's fma) and
counts mismatches against IEEE division over random and adversarial operands; the GPU
test test_device_division_by_reciprocal_correctly_rounded checks the device code itself."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_division_by_reciprocal_matches_ieee_division():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "div_rcp_check")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", exe,
                        os.path.join(HERE, "c", "div_rcp_check.c"), "-lm"], check=True)
        out = subprocess.run([exe, "20000000"], capture_output=True, text=True, check=True).stdout.split()
    cases, bad = int(out[0]), int(out[1])
    assert cases == 20000000
    assert bad == 0
