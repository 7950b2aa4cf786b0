"""Native (C) checks of arithmetic shortcuts used by the kernels.

const_div_check.c: the division-free x/3, x/6 of the FP64 stage kernel
(q = fma(x, z, x*z'), 1/d split into z = RN(1/d) and z' = RN(1/d - z)) equals
IEEE x/d bit for bit on random doubles with 2^-900 <= |x| <= 2^1000, integers,
quotients next to rounding boundaries
and signed zeros; and the WENO5 weight division through a reciprocal refined
by two Newton steps (started with a relative error up to 2^-22) equals IEEE a/b
over the operand ranges the weights take.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_constant_division_shortcut_is_exact(tmp_path):
    exe = str(tmp_path / "cdc")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "native", "const_div_check.c"),
                    "-lm"], check=True)
    r = subprocess.run([exe, "20000000"], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().splitlines()[-1] == "0"
