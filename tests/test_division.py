"""The flux divergence of the one-column-per-wavefront kernels divides a cell's edge fluxes by
its volume through one reciprocal and a Markstein correction (div_reuse); this must be
bit-identical to the IEEE quotient the reference computes.  The C checker evaluates the same
expression in the host's double precision (fused multiply-add from libm) against a / b."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_reciprocal_corrected_division_is_exact(tmp_path):
    exe = tmp_path / "div_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(HERE / "native" / "div_check.c"), "-o", str(exe),
                    "-lm"], check=True)
    out = subprocess.run([str(exe), "20000000"], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "0", out.stdout
