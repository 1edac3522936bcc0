"""Host test of the tile Cholesky's worker queue (nelson_amd/csrc/chol_tasks.h):
the early-queued pre-updates keep the task set and the column order of the
rest, and a simulation of the persistent launch (workgroups taking tasks in
order, each holding its task until the tiles and chain steps it reads are
complete) finishes every task at the look-ahead the library picks for every
launch width 1..255.  The checker is tests/cpp/chol_tasks_check.cc, compiled
here with g++ (no GPU)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_chol_task_queue(tmp_path):
    exe = str(tmp_path / "chol_tasks_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "nelson_amd", "csrc"),
                    os.path.join(HERE, "cpp", "chol_tasks_check.cc"), "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-2000:]
