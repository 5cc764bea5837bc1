"""CPU: the host-side pieces of bench.py — the HBM traffic lookup in the
committed counter summary, one design of the host process-pool sweep
baseline (the oracle's reference-arithmetic step with the design's source
and port DFTs), and the refusal of runs on several GPUs."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_traffic_lookup_reads_committed_summary():
    launch = {"vec": 1, "ty": 3, "kchunk": 12, "order": 3}
    got = bench.measured_traffic("voxel_3d_256_fp32_fp16", launch)
    assert got is not None
    # the counters put the fused step between the algorithmic bytes and
    # 1.5x them (writes exact, halo rows mostly L2 hits)
    alg = 1056964608.0
    assert alg <= got <= 1.5 * alg
    assert bench.measured_traffic("voxel_3d_256_fp32_fp16", dict(launch, kchunk=7)) is None
    assert bench.measured_traffic("no_such_workload", launch) is None


def test_cpu_sweep_leg_runs_one_design():
    secs = bench._cpu_coupler((0, 3))
    assert secs > 0


@pytest.mark.parametrize("n", [2, 8])
def test_several_gpus_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
