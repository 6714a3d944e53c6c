"""Host logic of the tree stage, on the CPU: the BP segment schedule and the
weight-order check of deedsbcv_amd/csrc/mst_host.cpp, compiled with g++ into
a checker (tests/host_schedule_check.cpp) that compares the schedule's
programs with the reference's depth-by-depth fold (regularisation.h:190-235)
on random trees, and the check with a brute force over distinct edge costs."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "deedsbcv_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_bp_schedule_and_weight_check(tmp_path):
    exe = tmp_path / "host_schedule_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(CSRC), str(ROOT / "tests" / "host_schedule_check.cpp"),
                    str(CSRC / "mst_host.cpp"), "-o", str(exe)], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout.splitlines()[0]
    assert "0 mismatches" in r.stdout.splitlines()[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_soft_sweep_schedule_on_random_trees(tmp_path):
    """The soft sweeps' schedule (re-rooting at the centre, rounds of both
    directions) checked by tests/soft_schedule_check.cpp against brute
    force on 1500 random trees."""
    exe = tmp_path / "soft_schedule_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(CSRC), str(ROOT / "tests" / "soft_schedule_check.cpp"),
                    str(CSRC / "mst_host.cpp"), "-o", str(exe)], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1].startswith("0 mismatches")
