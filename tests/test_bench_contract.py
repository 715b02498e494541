"""The benchmark's command-line contract, checked without a GPU.

bench.py measures one GPU: any other --gpus value ends with exit status 2 and
a message, before torch or the HIP library is loaded.  The measured HBM bytes
it reports beside the roofline come from profiles/<workload>_hbm_traffic.json;
it must hold the fluid step's dominant kernel, k_p2g, by the name the bench
looks up.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_more_than_one_gpu(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 2
    assert r.stdout == ""  # no JSON line for a run that measured nothing
    assert f"--gpus {n}" in r.stderr and "one GPU" in r.stderr


def test_bench_help_lists_contract_flags():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in r.stdout


def test_traffic_file_holds_the_dominant_kernel():
    with open(os.path.join(ROOT, "profiles", "blobs1024_hbm_traffic.json")) as f:
        t = json.load(f)
    assert t["workload"] == "blobs1024"
    per = t["bytes_per_launch"]
    assert per, "no kernels in the traffic file"
    assert "k_p2g" in {name.split("<")[0] for name in per}
    for name, b in per.items():
        assert isinstance(b, int) and b > 0, name
