"""Host tables of the scatter mode of HEX27 sets, checked on the CPU (no GPU).

Sets too large for a K_e pool assemble by colour classes, each element adding its K_e into the CSR
values at positions from per-element rank tables (moris_amd/csrc/set/set_host.cpp).  A small C++
program (tests/host/scatter_check.cpp) builds real sets through the host problem API, forces the
scatter mode and checks, for every element-matrix entry, that the computed position lies in the
entry's CSR row and holds the column of its adof, and that no two elements of one colour share an
adof (the condition that makes the plain read-modify-write stores of one launch race-free).

Sets that keep a K_e pool assemble by gather lists, one per CSR value (built per CSR row, rows in
parallel).  tests/host/gather_check.cpp checks the properties that fix those lists: every
element-matrix entry appears exactly once, in the list of the value at its row and column, and each
list ascends in the entry index (the order of summation)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "moris_amd", "lib")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    assert os.path.exists(os.path.join(LIBDIR, "libmoris_gpu.so")), "build the library first"
    exe = str(tmp_path_factory.mktemp("scatter") / "scatter_check")
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "moris_amd", "csrc", "set"), os.path.join(HERE, "host", "scatter_check.cpp"),
           os.path.join(ROOT, "moris_amd", "csrc", "set", "set_host.cpp"), "-L", LIBDIR, "-lmoris_gpu",
           "-Wl,-rpath," + LIBDIR, "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.parametrize("n,refine", [(6, 0), (5, 1), (8, 1)])
def test_scatter_mode_positions_and_colours(checker, n, refine):
    r = subprocess.run([checker, str(n), str(refine)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "OK:" in r.stdout, r.stdout + r.stderr


@pytest.fixture(scope="module")
def gather_checker(tmp_path_factory):
    assert os.path.exists(os.path.join(LIBDIR, "libmoris_gpu.so")), "build the library first"
    exe = str(tmp_path_factory.mktemp("gather") / "gather_check")
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "moris_amd", "csrc", "set"), os.path.join(HERE, "host", "gather_check.cpp"),
           os.path.join(ROOT, "moris_amd", "csrc", "set", "set_host.cpp"), "-L", LIBDIR, "-lmoris_gpu",
           "-Wl,-rpath," + LIBDIR, "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.parametrize("n,refine", [(6, 0), (5, 1), (8, 1)])
def test_gather_lists_cover_every_entry_once_in_order(gather_checker, n, refine):
    r = subprocess.run([gather_checker, str(n), str(refine)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "OK:" in r.stdout, r.stdout + r.stderr
