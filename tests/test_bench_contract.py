"""bench.py's output contract: one JSON line with the required keys, the
multi-GPU refusal, the CPU-baseline leg; and (GPU) a small full run whose legs
that do not fit the small buffers are skipped rather than overrun."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

REQUIRED = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
            "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline")


def _bench(*args, timeout=600):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=timeout, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    p = _bench("--gpus", "2")
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_parse_k():
    import bench
    assert bench.parse_k("1-16") == list(range(1, 17))
    assert bench.parse_k("5,16") == [5, 16]
    assert bench.parse_k("1-3,15") == [1, 2, 3, 15]


def test_cpu_baseline_leg_on_a_small_sample():
    import bench
    out = bench.cpu_baseline([1, 5], 12, 3)
    assert out["unit"] == "GB/s" and out["kind"] == "port" and out["value"] > 0 and out["cores"] >= 1
    assert "N=2^12" in out["sample"]


@pytest.mark.gpu
def test_small_bench_run_keeps_the_contract(cuda):
    p = _bench("--k", "1,5", "--n-log2", "18", "--steps", "2", "--warmup", "0", "--no-cpu", "--jobs", "64",
               "--requests", "256")
    assert p.returncode == 0, p.stderr[-3000:]
    lines = p.stdout.strip().splitlines()
    assert len(lines) == 1
    out = json.loads(lines[0])
    missing = [k for k in REQUIRED if k not in out]
    assert not missing, missing
    assert out["n_gpus"] == 1 and out["steps"] == 2 and out["value"] > 0 and out["higher_is_better"] is True
    r = out["roofline"]
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3
    assert out["block"] is None and "skipped" in p.stderr  # 2^18 structs: too small for the block leg
    assert out["batched"]["bit_exact"] and out["served"]["bit_exact_sampled"]
