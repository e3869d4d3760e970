"""Block order of the fused pass, checked on the CPU.

The pass can visit lattice blocks in linear or XCD-contiguous order, forward
or reversed (xadvcool_amd/csrc/lattice.h, launch_block).  The same header
function is compiled here with g++ and checked exhaustively to be a bijection
onto the lattice blocks with the XCD shares it promises, so an indexing
mistake is caught before any GPU run.  Planted indexing mistakes must fail
the check.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "block_order_check.cpp")


def _build_and_run(tmp_path, src):
    exe = tmp_path / "block_order_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), str(src)], check=True)
    return subprocess.run([str(exe)], capture_output=True, text=True)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_block_order_is_a_bijection_with_xcd_shares(tmp_path):
    r = _build_and_run(tmp_path, SRC)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("good,bad", [
    # the slot's offset left out of the position inside the XCD's eighth
    ("  const int lb = (b & 7) * per + (j - s * per);\n", "  const int lb = (b & 7) * per + j;\n"),
    # an eighth rounded down: the last blocks of a slot are never visited
    ("  const int per = (bps + 7) >> 3;\n", "  const int per = bps >> 3 > 0 ? bps >> 3 : 1;\n"),
])
def test_planted_block_order_mistakes_are_caught(tmp_path, good, bad):
    hdr = open(os.path.join(HERE, "..", "xadvcool_amd", "csrc", "lattice.h")).read()
    assert good in hdr
    (tmp_path / "lattice.h").write_text(hdr.replace(good, bad))
    src = open(SRC).read().replace('#include "../xadvcool_amd/csrc/lattice.h"', '#include "lattice.h"')
    (tmp_path / "check.cpp").write_text(src)
    r = _build_and_run(tmp_path, tmp_path / "check.cpp")
    assert r.returncode != 0, r.stdout
