"""Geometry of the LDS boundary kernel (nr-amssncku_amd/csrc/bssn_edge_tiles.h),
checked on the CPU by tests/edge_tiles_check.cpp built against the same
header: the stencil-order shells cover the array once with the order the
reference's room test gives each point (diff_new.f90:1084-1096), the tiles
give every shell point to one live thread, and every index a boundary
stencil may read -- within +-3 of the point, mirrored below a symmetry
plane (fmisc.f90:877-901) -- lies in the tile's LDS window."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "nr-amssncku_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_edge_tiles_cover_shells_and_stencil_reads(tmp_path):
    exe = str(tmp_path / "edge_tiles_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, os.path.join(HERE, "edge_tiles_check.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
