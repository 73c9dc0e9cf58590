"""Host-only checks of kernel arithmetic and geometry (no GPU): the K offset
tables' pad covers every index the gather's K loop reads, for every stage
depth the tile shapes give (both pipelines, every split of a split-K launch),
with the kernels' own constants (csrc/qs_tiles.h); and the INT8 CRT GEMM's
epilogue reduces INT32 sums mod m exactly (csrc/qs_crt.h)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_k_table_pad_covers_the_gather_reach(tmp_path):
    exe = str(tmp_path / "k_table_reach")
    src = os.path.join(ROOT, "tests", "native", "k_table_reach.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, src], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    # 96 instantiations: 2 element sizes x real/complex x 6 tiles x 4 gather directions.
    assert "configs 96" in r.stdout and "failures 0" in r.stdout, r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_crt_epilogue_reduction_is_exact(tmp_path):
    exe = str(tmp_path / "crt_mod_exact")
    src = os.path.join(ROOT, "tests", "native", "crt_mod_exact.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", exe, src], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "failures 0" in r.stdout, r.stdout
