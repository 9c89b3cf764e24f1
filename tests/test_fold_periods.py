"""Loop records of the exact multi-pool scheme (fold_periods in
mrilab_amd/csrc/decode.cpp), checked on the host: a C++ driver compiled
with g++ against decode.cpp folds synthetic record streams (a saturation
train with a symmetric pulse, rectangular rf, blocks holding ADC samples,
spoilers or captures, random streams) and checks that every stream unfolds
to its input and that only foldable blocks fold. The same driver checks
deferral_counts, which decides whether free precession is deferred, on a
stream with ADC samples, rf steps and a spoiler."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fold_periods_unfolds_to_its_input(tmp_path):
    exe = tmp_path / "fold_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "mrilab_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "fold_periods_check.cpp"),
                    os.path.join(ROOT, "mrilab_amd", "csrc", "decode.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
