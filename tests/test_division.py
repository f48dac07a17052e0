"""The Solver update divides by the neighbour count with a reciprocal product
and one residual correction instead of an IEEE divide (div_by_count and the
Solver kind of the sweep in taufactor_amd/csrc/tauhip_sweep.h); a subnormal
residual sends the quotient to the IEEE division.  This CPU test compiles
tests/csrc/divcheck.c, which replays that sequence with C's fmaf and compares
it bit for bit with the IEEE quotient for every divisor the Solver uses (a
zero of either sign counts as equal: the update is the same float).  The
default run covers every mantissa at the exponents around the subnormal-
residual boundary, the working range and the extremes;
TAUHIP_DIVCHECK_FULL=1 runs all 2^32 inputs (0 mismatches).  The GPU parity
tests cover the compiled kernel end to end."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_division_by_count_is_ieee(tmp_path):
    exe = tmp_path / "divcheck"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(HERE, "csrc", "divcheck.c"), "-o", str(exe), "-lm"], check=True)
    args = [str(exe)] + (["full"] if os.environ.get("TAUHIP_DIVCHECK_FULL") == "1" else [])
    res = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("mismatches=0") == 9, res.stdout
