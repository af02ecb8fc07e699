"""The short correctly rounded square root and divisions of the pair loops (dfsph_device.h: sqrt_rn, div_h, div_rn)
agree with IEEE sqrtf and division on every checked operand, modelled on the CPU with C's fmaf and hardware
estimates off by up to one ulp either way (tests/pair_math_check.c). On the GPU the same sequences are checked by every
bitwise comparison of a pair kernel with the oracle, whose sqrtf and divisions are IEEE's."""

from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_pair_arithmetic_is_correctly_rounded(tmp_path):
    exe = tmp_path / "pair_math_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), os.path.join(HERE, "pair_math_check.c"), "-lm"],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    for line in out.stdout.splitlines():
        if "checked" in line:
            assert line.endswith(" 0 wrong"), line
