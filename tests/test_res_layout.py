"""The resident solver's boundary-exchange layout (thirring2d_amd/csrc/
res_layout.h), checked on the host: a native check compiled with g++ from
tests/native/res_layout_check.cpp walks every even lattice up to 160^2."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "thirring2d_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_res_layout_covers_every_halo(tmp_path):
    exe = tmp_path / "res_layout_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC,
                    os.path.join(HERE, "native", "res_layout_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe), "160"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 bad" in out.stdout and " 0 not needed" in out.stdout, out.stdout
