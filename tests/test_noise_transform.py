"""The device generator's Box-Muller transform uses its own log and sin/cos
(series in sopg_device.h, the same arithmetic on host and device).  Compiled
here with g++ and checked against long-double libm over 4M arguments."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "sop-cpu_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_box_muller_log_and_sincos_within_few_ulp(tmp_path):
    exe = tmp_path / "bm_accuracy"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
                    os.path.join(HERE, "host", "bm_accuracy.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    ulp_log, ulp_sin, ulp_cos, abs_sc = map(float, out)
    assert ulp_log < 4.0 and ulp_sin < 4.0 and ulp_cos < 4.0, out
    assert abs_sc < 4e-16, out
