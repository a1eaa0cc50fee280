"""The exact closed form of the reference's sequential radiation ramp.

The panel and R4 radiation terms are applied by the reference as
t <- t + RDEL point after point (clblm_ODLAY.f90:2458-2481,
clblm_ODLAY_LineF4.f90:152-177).  The GPU evaluates that ramp in parallel from the
pieces of clblm_amd/csrc/seq_ramp.h; this test compiles the header for the CPU and
compares every point of 63 000 ramps (binade crossings up and down, exact rounding
ties, zero/tiny/large steps, radiation-like magnitudes) with the sequential loop,
bit for bit.  It also checks that the naive form t0 + j*d differs from the loop in
most of them, so the comparison can tell the two apart.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_ramp_pieces_bitwise():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "chk")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-o", exe,
                        os.path.join(HERE, "native", "seq_ramp_check.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    cases, bad, naive_bad = map(int, out[-3:])
    assert cases == 63000
    assert bad == 0
    assert naive_bad > cases // 2
