#!/usr/bin/env python3
"""Benchmark of the MI355X AoS load/store path.

Default workload (one "step"): the struct-size sweep -- structs of k 32-bit
words, k = 1..16 (4..64 B), N = 2^26 structs each, uniform random payload;
for every k
  * contiguous copy dst[i] = src[i] through load_warp_contiguous /
    store_warp_contiguous (the Trove low-level API), and
  * random gather dst[i] = src[idx[i]], idx a seeded uniform permutation of
    [0, N), through the tile path (indexed tile loads + coalesced stores) and
    through coalesced_ptr (the Trove high-level API).
The naive per-lane kernels run on the same inputs outside the timed steps for
the "vs naive" ratios. Inputs are resident in HBM before timing starts.

Algorithmic bytes (the reference harness's accounting, tests/benchmark.cu:
147, 210, 283): copy 2*S per struct, gather 2*S + 4 per struct.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n-log2", type=int, default=26, help="structs per struct size (log2)")
    p.add_argument("--k", type=str, default="1-16", help="struct sizes in 32-bit words, e.g. 1-16 or 5,16")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-n-log2", type=int, default=23, help="structs per size in the CPU sample (log2)")
    p.add_argument("--no-naive", action="store_true", help="skip the naive-kernel comparison")
    p.add_argument("--seed", type=int, default=20260101)
    return p.parse_args()


def parse_k(spec: str) -> list[int]:
    out: list[int] = []
    for part in spec.split(","):
        if "-" in part:
            a, b = part.split("-")
            out.extend(range(int(a), int(b) + 1))
        else:
            out.append(int(part))
    return out


class KernelClock:
    """HIP events around each launch on the launch stream (torch's current)."""

    def __init__(self, torch):
        self.torch = torch
        self.pending = []  # (name, bytes, start, end)
        self.totals = {}   # name -> [launches, ms, bytes]

    def run(self, name, nbytes, fn):
        t = self.torch
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        self.pending.append((name, nbytes, s, e))

    def collect(self):
        for name, nbytes, s, e in self.pending:
            rec = self.totals.setdefault(name, [0, 0.0, 0])
            rec[0] += 1
            rec[1] += s.elapsed_time(e)
            rec[2] += nbytes
        self.pending.clear()


def cpu_baseline(ks, n_log2, seed):
    """The oracle's OpenMP copy + gather on a bounded sample of the workload."""
    import numpy as np

    import oracle
    import trove_amd as T

    n = 1 << n_log2
    smax = 4 * max(ks)
    src = np.frombuffer(T.gen_words(n * smax // 4, seed).tobytes(), dtype=np.uint8)
    idx = T.gen_permutation(n, seed + 1)
    total_bytes, total_s = 0, 0.0
    for k in ks:
        s = 4 * k
        view = src[: n * s]
        oracle.copy(view, n, s, omp=True)          # warm-up (page faults)
        oracle.gather(view, idx, s, omp=True)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.copy(view, n, s, omp=True)
            oracle.gather(view, idx, s, omp=True)
            times.append(time.perf_counter() - t0)
        total_s += statistics.median(times)
        total_bytes += n * (2 * s) + n * (2 * s + 4)
    threads = oracle.lib().oracle_max_threads()
    return {
        "value": round(total_bytes / total_s / 1e9, 2),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle OpenMP copy + gather (memcpy per struct), k={ks[0]}..{ks[-1]} words, "
                   f"N=2^{n_log2} structs per size, median of 3, {threads} threads"),
    }


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        log("bench.py: only --gpus 1 is supported (one MI355X); see the sharding design notes")
        return 2

    import torch

    import trove_amd as T

    T.lib()  # fail loudly if the HIP library is missing
    assert torch.cuda.is_available(), "no GPU visible"
    dev = torch.device("cuda:0")
    ks = parse_k(args.k)
    n = 1 << args.n_log2
    smax = 4 * max(ks)
    log(f"bench: k={ks} N=2^{args.n_log2} steps={args.steps} warmup={args.warmup}")

    # Inputs resident in HBM: payload (device hash fill), permutation (host
    # Fisher-Yates, uploaded once).
    src = torch.empty(n * smax // 4, dtype=torch.int32, device=dev)
    dst = torch.empty_like(src)
    T.fill_words_device(src, src.numel(), args.seed)
    t0 = time.time()
    perm = T.gen_permutation(n, args.seed + 1)
    idx = torch.from_numpy(perm).to(dev)
    del perm
    torch.cuda.synchronize()
    log(f"bench: inputs ready ({time.time() - t0:.1f}s host permutation)")

    def step(clock):
        for k in ks:
            s = 4 * k
            clock.run(f"copy_shfl_k{k}", 2 * s * n, lambda: T.copy(src, dst, n, s, variant="shfl"))
            clock.run(f"gather_tile_k{k}", (2 * s + 4) * n, lambda: T.gather(src, idx, dst, n, s, variant="tile"))
            clock.run(f"gather_shfl_k{k}", (2 * s + 4) * n, lambda: T.gather(src, idx, dst, n, s, variant="shfl"))

    warm = KernelClock(torch)
    for _ in range(args.warmup):
        step(warm)
    torch.cuda.synchronize()
    warm.collect()

    clock = KernelClock(torch)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(clock)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    clock.collect()
    log(f"bench: timed {args.steps} steps in {wall:.3f}s")

    step_bytes = sum(v[2] for v in clock.totals.values()) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = step_bytes / (ms_per_step * 1e-3) / 1e9

    per = {}
    for name, (launches, ms, nbytes) in clock.totals.items():
        avg_ms = ms / launches
        per[name] = {"avg_ms": round(avg_ms, 4), "gbs": round(nbytes / launches / (avg_ms * 1e-3) / 1e9, 1)}
    dominant = max(clock.totals.items(), key=lambda kv: kv[1][1])
    dname, (dl, dms, dbytes) = dominant
    d_gbs = dbytes / dl / (dms / dl * 1e-3) / 1e9

    # Naive per-lane kernels and the plain-copy ceiling, same inputs, untimed
    # by the step clock.
    naive = {}
    if not args.no_naive:
        nclock = KernelClock(torch)
        for k in ks:
            s = 4 * k
            for _ in range(3):
                nclock.run(f"copy_direct_k{k}", 2 * s * n, lambda: T.copy(src, dst, n, s, variant="direct"))
                nclock.run(f"gather_direct_k{k}", (2 * s + 4) * n,
                           lambda: T.gather(src, idx, dst, n, s, variant="direct"))
        for _ in range(3):
            nclock.run("memcpy_ceiling", 2 * n * smax, lambda: T.memcpy_ceiling(src, dst, n * smax))
        torch.cuda.synchronize()
        nclock.collect()
        for name, (launches, ms, nbytes) in nclock.totals.items():
            avg_ms = ms / launches
            naive[name] = {"avg_ms": round(avg_ms, 4), "gbs": round(nbytes / launches / (avg_ms * 1e-3) / 1e9, 1)}

    sweep = []
    for k in ks:
        row = {"k": k, "bytes": 4 * k,
               "copy_gbs": per[f"copy_shfl_k{k}"]["gbs"],
               "gather_tile_gbs": per[f"gather_tile_k{k}"]["gbs"],
               "gather_coalesced_ptr_gbs": per[f"gather_shfl_k{k}"]["gbs"]}
        if naive:
            cd, gd = naive[f"copy_direct_k{k}"]["gbs"], naive[f"gather_direct_k{k}"]["gbs"]
            row.update({"copy_naive_gbs": cd, "gather_naive_gbs": gd,
                        "copy_vs_naive": round(row["copy_gbs"] / cd, 2),
                        "gather_tile_vs_naive": round(row["gather_tile_gbs"] / gd, 2),
                        "gather_coalesced_ptr_vs_naive": round(row["gather_coalesced_ptr_gbs"] / gd, 2)})
        row["copy_frac_peak"] = round(row["copy_gbs"] / HBM_PEAK_GBS, 3)
        sweep.append(row)

    traffic = None
    traffic_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if os.path.exists(traffic_path):
        with open(traffic_path) as f:
            tr = json.load(f)
        if dname in tr:
            traffic = tr[dname]

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, OpenMP)")
        cpu = cpu_baseline(ks, args.cpu_n_log2, args.seed)

    out = {
        "metric": "AoS copy/gather GB/s (% HBM peak) vs naive loads",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded hash payload, Fisher-Yates permutation)",
        "config": {
            "workload": (f"struct-size sweep: k={ks[0]}..{ks[-1]} x 32-bit words, N=2^{args.n_log2} structs per "
                         "size; contiguous copy via load/store_warp_contiguous + random-permutation gather via "
                         "tile path and coalesced_ptr"),
            "structs_per_size": n,
            "struct_bytes": [4 * k for k in ks],
            "parallelism": "single GPU",
        },
        "frac_peak": round(value / HBM_PEAK_GBS, 3),
        "roofline": {
            "bound": "hbm",
            "kernel": dname,
            "achieved": round(d_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(d_gbs / HBM_PEAK_GBS, 3),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "memcpy_ceiling_gbs": naive.get("memcpy_ceiling", {}).get("gbs"),
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
