"""Stage regions of the batched forms on a block hierarchy (host code).

A block that reads some guard sectors from same-level leaves computes its
interior and the other sectors at each stage's telescoping depth; the
region is cut into boxes (flash-x_amd/csrc/spark_regions.hpp).  The checker
compiles against that header with the host C++ compiler and tests, for the
64 unions of whole faces and 2000 pseudo-random sector masks at depths
6 / 3 / 0, that the boxes are disjoint and cover exactly that set.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_sector_boxes_cover_the_computed_sectors_exactly(tmp_path):
    exe = tmp_path / "check_sector_boxes"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", str(ROOT / "flash-x_amd" / "csrc"),
                    str(ROOT / "tests" / "cpp" / "check_sector_boxes.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip())
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_boundary_cells_visit_each_boundary_cell_once(tmp_path):
    """k_correct's thread-per-boundary-cell enumeration (spark_regions.hpp):
    every cell with index 0 or n-1 in an active dimension exactly once, no
    interior cell, for 1-3 dimensions and extents 1-6 and 16."""
    exe = tmp_path / "check_boundary_cells"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", str(ROOT / "flash-x_amd" / "csrc"),
                    str(ROOT / "tests" / "cpp" / "check_boundary_cells.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip())
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout
