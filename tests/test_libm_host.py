"""The device's libm restatements (csrc/rt_libm.h) built for the host and compared bit
for bit with the host C library -- glibc's exp, log and log10, the functions the
reference's EXP, LOG and LOG10 call -- on 10 M arguments plus special values."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from conftest import PKG, ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_libm_restatement_is_the_host_libm(tmp_path):
    exe = str(tmp_path / "libm_host_check")
    subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-I", os.path.join(PKG, "csrc"),
                    os.path.join(ROOT, "tests", "libm_host_check.cc"), "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe, "10000000"], capture_output=True, text=True)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert " exp 0, log 0, log10 0 differ" in r.stdout
