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


def _counter_csv(path, counter, dispatches):
    # dispatches: (kernel name, grid size, KiB) in dispatch order
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Grid_Size,Counter_Name,Counter_Value\n")
        for d, (name, grid, kib) in enumerate(dispatches, start=1):
            # Counters of several hardware instances appear as several rows.
            f.write(f'{d},"{name}",{grid},{counter},{kib / 2}\n')
            f.write(f'{d},"{name}",{grid},{counter},{kib / 2}\n')


def test_traffic_from_attributes_dispatches_to_legs(tmp_path):
    import bench

    M = ("void trove_amd::k_memcpy16(uint4 const*, uint4*, long)", 256)
    legs = [["copy_a", 2000], ["memcpy_ceiling", 8192], ["gather_b", 1000], ["copy_a", 2000]]
    fetch = [("void trove_amd::k_fill_words(unsigned int*, long, unsigned long)", 4096, 99.0),  # set-up
             (*M, 0.0), ("void trove_amd::k_copy<x>()", 1024, 1.0), ("void trove_amd::k_copy_tail<x>()", 64, 0.5),
             (*M, 0.0), ("void trove_amd::k_memcpy16(uint4 const*, uint4*, long)", 8192, 2.0),
             (*M, 0.0), ("at::native::elementwise_kernel<...>", 4096, 50.0), ("void trove_amd::k_gather<x>()", 512, 3.0),
             (*M, 0.0), ("void trove_amd::k_copy<x>()", 1024, 2.5)]
    write = [(n, g, 2 * k) for n, g, k in fetch]
    _counter_csv(str(tmp_path / "f" / "host" / "1_counter_collection.csv"), "FETCH_SIZE", fetch)
    _counter_csv(str(tmp_path / "w" / "host" / "1_counter_collection.csv"), "WRITE_SIZE", write)
    with open(tmp_path / "legs.json", "w") as f:
        json.dump({"legs": legs}, f)
    out = tmp_path / "traffic.json"
    assert bench.traffic_from(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "legs.json"), str(out)) == 0
    tr = json.load(open(out))
    # memcpy_ceiling: 4096 bytes each way read 2 KiB fetched and 4 KiB written.
    assert tr["calibration"]["fetch_factor"] == 2.0 and tr["calibration"]["write_factor"] == 1.0
    k = tr["kernels"]
    # copy_a: legs of 1.5 and 2.5 KiB fetched (the tail kernel counts), twice that written.
    assert k["copy_a"] == round(2.0 * 1024 * 2 + 4.0 * 1024 * 1)
    assert k["gather_b"] == round(3 * 1024 * 2 + 6 * 1024)  # torch's kernel is not the leg's
    assert k["memcpy_ceiling"] == 8192
    assert tr["kernel_sources_sha256"] == bench.kernel_sources_sha256()
    kernels, stale = bench.load_traffic(str(out))
    assert not stale and kernels == k
    split, _ = bench.load_traffic(str(out), split=True)
    assert split["gather_b"] == (3 * 1024 * 2, 6 * 1024) and split["memcpy_ceiling"] == (4096, 4096)
    # The HBM bound prices random row fills (reads beyond the index stream) at
    # the random-line rate and the rest at the streaming rate.
    assert bench.hbm_bound_ms(10e9, 4e9, 1e9, True, 5.0, 6.0) == pytest.approx((9 / 5 + 5 / 6) * 1e3)
    assert bench.hbm_bound_ms(4e9, 4e9, 0, False, 5.0, 8.0) == pytest.approx(1e3)
    tr["kernel_sources_sha256"] = "0" * 64  # measured on other sources: not used
    json.dump(tr, open(out, "w"))
    assert bench.load_traffic(str(out)) == (None, True)
    # The generic per-leg counter table reads the same runs.
    import contextlib
    import io

    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        assert bench.counters_from([str(tmp_path / "f"), str(tmp_path / "w")], str(tmp_path / "legs.json")) == 0
    lines = buf.getvalue().splitlines()
    assert lines[0] == "| leg | FETCH_SIZE | WRITE_SIZE |"
    assert "| `copy_a` | 2 | 4 |" in lines and "| `gather_b` | 3 | 6 |" in lines
    # A leg list that does not match the marked dispatches is refused.
    with open(tmp_path / "legs.json", "w") as f:
        json.dump({"legs": legs[:3]}, f)
    with pytest.raises(SystemExit):
        bench.traffic_from(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "legs.json"), str(out))


def test_trace_from_summarises_kernel_traces(tmp_path):
    import contextlib
    import io

    import bench

    d = tmp_path / "t" / "host"
    os.makedirs(d)
    g15 = "void trove_amd::k_gather_shfl<trove::array<unsigned int, 15> >(trove::array<unsigned int, 15> const*, int const*, trove::array<unsigned int, 15>*, long)"
    rows = [(g15, 67108864, 0, 3000000), (g15, 67108864, 4000000, 7000000), (g15, 4096, 8000000, 8001000),
            ("void trove_amd::k_batch_gather(trove_amd::job_desc const*, int, long, int*)", 8192, 20000000, 20100000),
            ("void trove_amd::k_batch_gather_pack(trove_amd::job_pack)", 4096, 20110000, 20150000)]
    with open(d / "1_kernel_trace.csv", "w") as f:
        f.write("Kernel_Name,Grid_Size_X,Grid_Size_Y,Grid_Size_Z,Start_Timestamp,End_Timestamp\n")
        for name, grid, t0, t1 in rows:
            f.write(f'"{name}",{grid},1,1,{t0},{t1}\n')
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        assert bench.trace_from(str(tmp_path / "t"), None, 26) == 0
    out = buf.getvalue().splitlines()
    # Per kernel and grid; the sweep table takes the largest grid: 2^26 rows
    # of 124 algorithmic bytes in 3,000 us.
    assert "| `k_gather_shfl<array<u32, 15> >` | 67108864 | 2 | 3000.0 | 3000.0 | 3000.0 | 6.00 | 97.7% |" in out
    assert any(line.startswith("| `k_gather_shfl` | 15 | 67108864 | 2 | 3000.0 | 2774 |") for line in out)
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        assert bench.trace_from(str(tmp_path / "t"), "k_batch", 26) == 0
    out = buf.getvalue().splitlines()
    assert out[2] == "| 0 | 0 | `k_batch_gather` | 8192 | 100.0 |  |"
    assert out[3] == "| 0 | 1 | `k_batch_gather_pack` | 4096 | 40.0 | 10.0 |"


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
