"""The benchmark's output contract: one JSON line with the metric, the roofline of the dominant
kernel (achieved = algorithmic bytes per tet x tets / kernel time, frac = achieved / peak)
and the CPU baselines; more than one GPU is refused with a message before any GPU work."""

import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120, cwd=ROOT)
    assert res.returncode == 2
    line = json.loads(res.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


@pytest.mark.gpu
def test_bench_line_small_workload():
    print("[bench contract] n=8, short run", flush=True)
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--n", "8", "--steps", "20", "--warmup", "5",
           "--no-elasticity", "--no-md", "--batch-solve-instances", "4", "--batch-solve-nmax", "5",
           "--cpu-seconds", "1", "--cpu-solve-n", "0"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert res.returncode == 0, res.stderr[-2000:]
    d = json.loads(res.stdout.strip().splitlines()[-1])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 20 and d["warmup"] == 5 and d["value"] > 0
    E = d["config"]["elements"]
    assert E == 8 ** 3 * 6
    rl = d["roofline"]
    achieved = rl["alg_bytes_per_tet"] * E / (rl["kernel_ms"] * 1e-3) / 1e9
    assert abs(achieved - rl["achieved"]) <= 0.02 * achieved
    assert abs(rl["frac"] - rl["achieved"] / rl["peak"]) <= 1e-3
    cpu = d["cpu_baseline"]
    assert cpu["kind"] == "port" and cpu["cores"] == 1 and cpu["value"] > 0
    ac = cpu["all_core"]
    assert ac["cores"] >= 1 and ac["solve_residual"] <= 1e-8 and ac["value"] > 0
    bs = d["batch_solve"]
    assert bs["converged"] == 4 and bs["twogrid_converged"] == 4 and bs["learned_converged"] == 4
