"""bench.py pieces that run without a GPU: the one-GPU guard, the traffic lookup in
the committed rocprofv3 summary, and the CPU baselines on small cases."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from pynof_amd import synth, workcount  # noqa: E402


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("table", ["SPLIT_KERNEL", "FP64_KERNEL"])
def test_dominant_kernels_have_profiled_traffic(table):
    """The kernel names bench.py looks up for `roofline.traffic` are rows of the committed
    summary (a tile change renames the kernel; the summary and the names must move together)."""
    names = getattr(bench, table)
    tr = bench.profiled_traffic(names["half_transform"])
    assert tr is not None, names["half_transform"]
    assert tr["bytes"] > 1e9 and tr["fetch_GiB"] > 0 and tr["write_GiB"] > 0


def test_profiled_traffic_unknown_kernel():
    assert bench.profiled_traffic("k_no_such_kernel<1>") is None


def test_cpu_baseline_small():
    w = synth.workload(0)
    r = bench.cpu_baseline(w, w.naux)
    assert r["kind"] == "port" and r["unit"] == "builds/s" and r["value"] > 0
    assert f"naux={w.naux} of {w.naux}" in r["sample"]


def test_cpu_baseline_batch_sample():
    ws = synth.batch_workloads(8, lo=24, hi=48)
    flops = sum(workcount.build_flops(w.nbf, w.naux, w.nbf5) for w in ws)
    r = bench.cpu_baseline_batch(ws, flops, every=4, naux_div=4)
    assert r["value"] > 0 and "2 of 8 instances" in r["sample"]
