"""bench.py's contract on a machine without a GPU: more than one GPU is refused with
a message and exit status 2 for every workload, before anything touches a device;
the CPU-baseline leg (the oracle's RK4, timed) reports the documented fields."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT


@pytest.mark.parametrize("workload", ["headline", "ensemble", "precision", "adjoint", "variants"])
def test_more_than_one_gpu_is_refused(workload):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--workload", workload],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_cpu_baseline_fields():
    sys.path.insert(0, ROOT)
    import bench
    b = bench.cpu_baseline(16, 2)
    assert b["unit"] == "GSite*RK4-step/s" and b["kind"] == "port" and b["cores"] == 1
    assert b["value"] > 0 and b["all_cores"]["value"] > 0 and b["all_cores"]["cores"] >= 1
    assert "N=16^3" in b["sample"] and "2 step(s)" in b["sample"]
