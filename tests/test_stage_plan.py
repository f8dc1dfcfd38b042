"""The staged walk's copy plan, checked on the host: tests/stage_plan_check.cpp runs every
stage's segment copies on a byte image of the surpluses and checks that each staged leaf
finds its own bytes at its LDS offset, that the waves' copies never overlap, that they
stay inside the stage's buffer, and that full segments come before partial ones."""
import os
import re
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
    # the Python mirror of the sparse-order rule agrees with the host planner's
    from adaptive_sparse_grid_bench_amd import sgrid
    seen = 0
    for m in re.finditer(r"d=(\d+) n=\d+ elem=(\d) block-staged \((sparse|depth-first) order\)",
                         r.stdout):
        d, prec = int(m.group(1)), "f64" if m.group(2) == "8" else "f32"
        assert sgrid.blk_sparse(d, prec) == (m.group(3) == "sparse"), m.group(0)
        seen += 1
    assert seen >= 8
