"""CPU checks of the coarse-level elimination schedule (csrc/coarse_schedule.cpp).

tests/native/coarse_schedule_check.cpp runs the workgroup tasks of every launch
in plain double arithmetic on a random SPD matrix with the structure the
schedule assumes, and checks that the tasks of one launch never touch a tile
another task of that launch writes, that L L^T equals the permuted matrix and
that W (built from the schedule's row lists) satisfies W L = I.  No GPU call
is made."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "native", "coarse_schedule_check.cpp"),
       os.path.join(ROOT, "topopt-in-openmp-gpu_amd", "csrc", "coarse_schedule.cpp")]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("sched") / "coarse_schedule_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-o", exe] + SRC, check=True)
    return exe


@pytest.mark.parametrize("n,bw,planes", [
    (17 * 27, 0, 17),    # 17 planes of one tile (3x3 nodes per plane)
    (13 * 75, 0, 13),    # 13 planes of two tiles (5x5 nodes)
    (3 * 243, 0, 3),     # 3 planes of 9x9 nodes (4 tiles, 13 padding DOFs)
    (2 * 48, 0, 2),      # 2 planes
    (700, 150, 0),       # banded, ragged last tile
    (64, 63, 0),         # one tile
])
def test_schedule_factorizes_and_inverts(checker, n, bw, planes):
    out = subprocess.run([checker, str(n), str(bw), str(planes), "7"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK"), out.stdout
