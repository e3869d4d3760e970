"""Compact fluid storage: the neighbour numbering of the HIP pass, checked on the CPU.

The fused pass stores the fluid lattice for non-solid cells only and finds a
neighbour's slot from per-column tables (xadvcool_amd/csrc/lattice.h).  The
same header functions are compiled here with g++ and compared with a
brute-force ranking of non-solid cells on random masks, so an indexing
mistake is caught before any GPU run.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_compact_numbering_matches_brute_force(tmp_path):
    exe = tmp_path / "compact_index_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), os.path.join(HERE, "compact_index_check.cpp")],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout
