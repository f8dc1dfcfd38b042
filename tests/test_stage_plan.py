"""The staged walk's copy plan, checked on the host: tests/stage_plan_check.cpp runs every
stage's segment copies on a byte image of the surpluses and checks that each staged leaf
finds its own bytes at its LDS offset, that the waves' copies never overlap, that they
stay inside the stage's buffer, and that full segments come before partial ones."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(ROOT, "adaptive_sparse_grid_bench_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_stage_plans_copy_every_leaf(tmp_path):
    exe = str(tmp_path / "stage_plan_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC,
                    os.path.join(ROOT, "tests", "stage_plan_check.cpp"),
                    os.path.join(CSRC, "sg_grid.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ALL OK")
    assert "d=10 n=8 elem=8" in r.stdout
