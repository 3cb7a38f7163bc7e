"""The differentiable path's host-side particle order (cpic_amd/csrc/
host_sort.h, split over host threads) against a sequential restatement:
identical permutations, built and run as a host-only program."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_threaded_sort_matches_sequential(tmp_path):
    exe = tmp_path / "host_sort_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-Wall", "-Wextra",
                    "-I", os.path.join(ROOT, "cpic_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host_sort_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "DIFFERENT" not in r.stdout and "round trip: ok" in r.stdout
