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
the "vs naive" ratios, twice: as the reference writes them (one struct per
thread, ordinary accesses, a grid-stride loop) and in the library's own
launch (DIRECT_LAUNCH: the same grid, cache policies and groups in flight),
which isolates what the transposition adds. Inputs are resident in HBM
before timing starts.

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
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n-log2", type=int, default=26, help="structs per struct size (log2)")
    p.add_argument("--k", type=str, default="1-16", help="struct sizes in 32-bit words, e.g. 1-16 or 5,16")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-n-log2", type=int, default=26,
                   help="structs per size in the CPU sample (log2; 26 = the sweep's N, about 10-15 s of CPU work)")
    p.add_argument("--no-naive", action="store_true", help="skip the naive-kernel comparison")
    p.add_argument("--seed", type=int, default=20260101)
    p.add_argument("--config", choices=("all", "sweep", "batched", "served", "line-order"), default="all",
                   help="all: the sweep line with batched and served summaries attached; line-order: the "
                        "random-line read rate with the line indices sorted within windows")
    p.add_argument("--jobs", type=int, default=1024, help="batched: number of gather jobs")
    p.add_argument("--tune-table", type=str, default=None, help="batched: tuner table to load instead of tuning")
    p.add_argument("--serve-slots", type=int, default=1024, help="served: request slots of the server")
    p.add_argument("--serve-variant", choices=("tile", "lds", "shfl", "direct", "auto"), default="tile",
                   help="served: gather path of the server's steps")
    p.add_argument("--requests", type=int, default=8192, help="served: number of requests")
    p.add_argument("--profile-legs", type=str, default=None, metavar="PATH",
                   help="profiling run: every leg launches once after a one-block marker copy and a "
                        "synchronisation; the leg order is written to PATH (JSON) for --traffic-from")
    p.add_argument("--traffic-from", nargs=3, default=None, metavar=("FETCH_DIR", "WRITE_DIR", "LEGS_JSON"),
                   help="no GPU: build profiles/traffic.json from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                        "output directories of a --profile-legs run, print the table as markdown")
    p.add_argument("--trace-from", default=None, metavar="DIR",
                   help="no GPU: summarise the rocprofv3 --kernel-trace output under DIR as markdown: per "
                        "kernel and grid size, and the sweep kernels' algorithmic GB/s at N = 2^--n-log2")
    p.add_argument("--timeline", default=None, metavar="PREFIX",
                   help="with --trace-from: instead, the dispatches of kernels whose name starts with PREFIX "
                        "in time order with their durations and the gaps between them")
    p.add_argument("--counters-from", nargs="+", default=None, metavar="DIR",
                   help="no GPU: DIR... LEGS_JSON: print every counter of rocprofv3 --pmc runs of a "
                        "--profile-legs run per leg (mean over the leg's launches) as a markdown table")
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


ROOT = os.path.dirname(os.path.abspath(__file__))
TRAFFIC_PATH = os.path.join(ROOT, "profiles", "traffic.json")


# The translation units whose kernels profiles/traffic.json measures (copy,
# gather, scatter, store / load, block interface, ceilings, permutation
# inversion) and the headers they include; the batch engine and the server
# (batch.hip, batch.h, serve.hip) have no leg in that table, and the C ABI
# header (include/trove_amd.h) holds no device code.
TRAFFIC_SOURCES = ("common.h", "gather_body.h", "copy.hip", "gather.hip", "block.hip", "capi.hip")


def kernel_sources_sha256() -> str:
    """Hash of the sources of the kernels profiles/traffic.json measures
    (TRAFFIC_SOURCES under trove_amd/csrc, and the device headers under
    include/trove): the stamp that ties the table to the kernels it was
    measured on."""
    import hashlib

    h = hashlib.sha256()
    files = [os.path.join(ROOT, "trove_amd", "csrc", f) for f in TRAFFIC_SOURCES]
    for d, _, names in os.walk(os.path.join(ROOT, "include", "trove")):
        files += [os.path.join(d, f) for f in names if f.endswith((".hip", ".h", ".cpp"))]
    for f in sorted(files, key=lambda x: os.path.relpath(x, ROOT)):
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


class ProfileLegs:
    """--profile-legs: the leg sequence of a profiling run. Each leg is one
    launch preceded by a marker (a one-block trove_memcpy_ceiling of 16
    bytes, a k_memcpy16 dispatch of grid size 256) and a synchronisation, so
    that the dispatches of a counter run can be attributed to legs in order."""

    legs = None  # list of [name, algorithmic bytes] when profiling
    _marker = None

    @classmethod
    def run(cls, torch, T, name, nbytes, fn):
        if cls._marker is None:
            cls._marker = torch.zeros(8, dtype=torch.int32, device="cuda:0")
        torch.cuda.synchronize()
        T.memcpy_ceiling(cls._marker[:4], cls._marker[4:], 16)
        fn()
        torch.cuda.synchronize()
        cls.legs.append([name, int(nbytes)])


class KernelClock:
    """HIP events around each launch on the launch stream (torch's current).
    The first launch of every name is untimed: HIP loads a kernel's code
    object lazily at its first launch. In a --profile-legs run every launch
    is a marked leg instead (no timing)."""

    def __init__(self, torch, seen=None):
        self.torch = torch
        self.pending = []  # (name, bytes, start, end)
        self.totals = {}   # name -> [launches, ms, bytes]
        self.samples = {}  # name -> [ms per launch]
        self.seen = set() if seen is None else seen

    def run(self, name, nbytes, fn):
        t = self.torch
        if ProfileLegs.legs is not None:
            import trove_amd as T

            ProfileLegs.run(t, T, name, nbytes, fn)
            return
        if name not in self.seen:
            self.seen.add(name)
            fn()
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        self.pending.append((name, nbytes, s, e))

    def collect(self):
        if ProfileLegs.legs is not None:
            return
        for name, nbytes, s, e in self.pending:
            rec = self.totals.setdefault(name, [0, 0.0, 0])
            ms = s.elapsed_time(e)
            rec[0] += 1
            rec[1] += ms
            rec[2] += nbytes
            self.samples.setdefault(name, []).append(ms)
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
        for _ in range(5):
            t0 = time.perf_counter()
            oracle.copy(view, n, s, omp=True)
            oracle.gather(view, idx, s, omp=True)
            times.append(time.perf_counter() - t0)
        total_s += statistics.median(times)
        total_bytes += n * (2 * s) + n * (2 * s + 4)
    threads = oracle.lib().oracle_max_threads()
    return {
        "value": float(f"{total_bytes / total_s / 1e9:.4g}"),  # 4 significant digits, also for tiny samples
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle OpenMP copy + gather (memcpy per struct), k={ks[0]}..{ks[-1]} words, "
                   f"N=2^{n_log2} structs per size, median of 5 after a warm-up, {threads} OpenMP threads "
                   f"(nproc {os.cpu_count()})"),
    }


def timed_runs(torch, fn, reps):
    out = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        out.append(e0.elapsed_time(e1))
    return out


def bench_batched(args, torch, T):
    """1024 independent gather jobs: struct size uniform over 4..64 B, length
    log-uniform 2^10..2^22 structs, index pattern uniform over identity /
    uniform random permutation / sorted runs (geometric, mean 32); source
    length = job length. Persistent launch vs its hipGraph replay vs one
    launch per job, at equal output (bit-exact compared)."""
    import numpy as np

    from trove_amd import tune
    from trove_amd.workloads import BatchedWorkload

    W = BatchedWorkload(args.jobs, args.seed, torch)
    nj = W.n_jobs
    log(f"bench batched: {nj} jobs, {W.structs} structs, inputs in {W.setup_s:.1f}s")
    if args.tune_table and os.path.exists(args.tune_table):
        table = tune.load(args.tune_table)
    else:
        t0 = time.time()
        table = tune.autotune(reps=3, seed=args.seed)
        log(f"bench batched: tuned {len(table['entries'])} shapes in {time.time() - t0:.1f}s")
    choice = {}

    def choose(s, pattern, L):
        v, c = tune.choose(table, s, pattern, L)
        choice[v] = choice.get(v, 0) + 1
        return v, c

    ba = T.Batch(W.jobs(choose, W.dst_a))
    bb = T.Batch(W.jobs(lambda s, p, L: tune.choose(table, s, p, L), W.dst_b))
    items, nbytes = ba.stats()
    ba.run("launch")
    bb.run("looped")
    torch.cuda.synchronize()
    dst_a, dst_b = W.dst_a, W.dst_b
    same = all(torch.equal(dst_a[i:i + (1 << 26)], dst_b[i:i + (1 << 26)]) for i in range(0, dst_a.numel(), 1 << 26))
    reps = 5
    t_pers = statistics.median(timed_runs(torch, lambda: ba.run("launch"), reps))
    ba.run("graph")
    t_graph = statistics.median(timed_runs(torch, lambda: ba.run("graph"), reps))
    t_loop = statistics.median(timed_runs(torch, lambda: bb.run("looped"), reps))
    # The resident-grid (persistent) kernel, into dst_b, checked against the
    # one-wavefront-per-item launch.
    bb.run("resident")
    torch.cuda.synchronize()
    same = same and all(torch.equal(dst_a[i:i + (1 << 26)], dst_b[i:i + (1 << 26)])
                        for i in range(0, dst_a.numel(), 1 << 26))
    t_res = statistics.median(timed_runs(torch, lambda: bb.run("resident"), reps))
    bound = batched_bound(torch, T, W, args.seed)
    # Spot-check a few jobs against the oracle on the host.
    import oracle

    checked = 0
    for j in list(range(0, nj, max(1, nj // 8)))[:8]:
        w0, nw = W.job_words(j)
        s = int(W.sizes[j])
        src_h = W.src[w0:w0 + nw].cpu().numpy().view(np.uint8)
        want = oracle.gather(src_h, W.job_indices(j), s)
        got = dst_a[w0:w0 + nw].cpu().numpy().view(np.uint8)
        same = same and np.array_equal(got, want)
        checked += 1
    ba.close()
    bb.close()
    structs = W.structs
    del W
    return {
        "workload": (f"{nj} gather jobs, struct 4..64 B uniform, length log-uniform 2^10..2^22, pattern uniform "
                     "{identity, random permutation, sorted runs mean 32}"),
        "jobs": nj, "structs": structs, "work_items": items, "algorithmic_bytes": nbytes,
        "modes": {"one_launch": "one launch, one wavefront per work item (TROVE_BATCH_LAUNCH)",
                  "graph": "that launch replayed from a hipGraph (TROVE_BATCH_GRAPH)",
                  "resident": "one launch of only the resident wavefronts, each striding over the work items "
                              "(TROVE_BATCH_RESIDENT, a persistent kernel)",
                  "looped": "one trove_gather launch per job (TROVE_BATCH_LOOPED)"},
        "one_launch_ms": round(t_pers, 3), "graph_ms": round(t_graph, 3), "resident_ms": round(t_res, 3),
        "looped_ms": round(t_loop, 3), "bound_ms": bound["ms"], "bound": bound["what"],
        "one_launch_jobs_per_s": round(nj / (t_pers * 1e-3), 1),
        "graph_jobs_per_s": round(nj / (t_graph * 1e-3), 1),
        "resident_jobs_per_s": round(nj / (t_res * 1e-3), 1),
        "looped_jobs_per_s": round(nj / (t_loop * 1e-3), 1),
        "one_launch_gbs": round(nbytes / (t_pers * 1e-3) / 1e9, 1),
        "graph_gbs": round(nbytes / (t_graph * 1e-3) / 1e9, 1),
        "resident_gbs": round(nbytes / (t_res * 1e-3) / 1e9, 1),
        "speedup_one_launch_vs_looped": round(t_loop / t_pers, 3),
        "speedup_graph_vs_looped": round(t_loop / t_graph, 3),
        "speedup_resident_vs_looped": round(t_loop / t_res, 3),
        "speedup_bound_vs_looped": round(t_loop / bound["ms"], 3),
        "variants_chosen": choice,
        "bit_exact": bool(same), "oracle_checked_jobs": checked,
    }


def batched_bound(torch, T, W, seed):
    """The batch priced job by job at standalone rates: every (struct size,
    index pattern) is timed as one trove_gather of 2^22 structs (the largest
    job length) through each gather path, and each job costs its algorithmic
    bytes at the best of those rates. Their sum is what one launch would
    take if each job ran at its pattern's large-job rate."""
    L = 1 << 22
    dev = torch.device("cuda:0")
    src = torch.empty(L * 16, dtype=torch.int32, device=dev)
    dst = torch.empty_like(src)
    T.fill_words_device(src, src.numel(), seed + 51)
    rate = {}
    for pattern in ("identity", "random", "runs"):
        idx = torch.from_numpy(T.gen_indices(L, L, pattern, seed + 52)).to(dev)
        for s in range(4, 65, 4):
            best = None
            for v in ("tile", "lds", "shfl", "direct"):
                T.gather(src, idx, dst, L, s, variant=v)  # first launch loads the code object
                ms = statistics.median(timed_runs(torch, lambda: T.gather(src, idx, dst, L, s, variant=v), 3))
                best = ms if best is None else min(best, ms)
            rate[(s, pattern)] = (2 * s + 4) * L / (best * 1e-3)
        del idx
    ms = sum((2 * int(s) + 4) * int(n) / rate[(int(s), p)] for s, n, p in zip(W.sizes, W.lengths, W.patterns)) * 1e3
    del src, dst
    return {"ms": round(ms, 3),
            "what": "sum over jobs of algorithmic bytes / best standalone rate of the job's (size, pattern) "
                    "at 2^22 structs over tile, LDS, coalesced_ptr and naive paths"}


def bench_served(args, torch, T):
    """Served gather: table of 2^24 rows x 32 B; requests of log-uniform
    2^8..2^18 indices drawn Zipf(1.1) over the table, indices device-resident;
    a saturation run (all requests queued at t = 0) fixes the request rate,
    then Poisson arrivals at 25%, 50%, 80% and 95% of it; 4 GiB output
    budget."""
    import numpy as np

    from trove_amd.workloads import ServedWorkload

    W = ServedWorkload(args.requests, args.seed, torch)
    table, n_rows, rb = W.table, W.n_rows, W.row_bytes
    R, sizes, offs, host_idx = W.n_requests, W.sizes, W.offs, W.host_idx
    torch.cuda.synchronize()
    log(f"bench served: {R} requests, {host_idx.size} indices, inputs in {W.setup_s:.1f}s")
    ptrs = W.device_pointers()
    budget = 4 << 30
    total_bytes = 68 * int(sizes.sum())

    # Slices of 2^18 indices, the workload's largest request, so that every
    # request finishes in one step (measured: 2^15-2^17 give 2-12% lower
    # saturation and up to 1.5x the p99 at 95% load).
    slice_indices = 1 << 18

    marker = torch.zeros(8, dtype=torch.int32, device=table.device)

    def mark_run():
        # A one-block marker copy before each run (outside its timing), so
        # that a kernel trace of this config splits into its runs.
        T.memcpy_ceiling(marker[:4], marker[4:], 16)
        torch.cuda.synchronize()

    # The same requests' indices in host memory (each submission range-checks
    # and copies its indices; the server uploads them before the step).
    host_ptrs = [host_idx.ctypes.data + 4 * int(o) for o in offs]

    def run(arrivals, slice_indices=slice_indices, on_device=True):
        k = len(arrivals)
        srv = T.Server(table, n_rows, rb, budget_bytes=budget, max_slots=args.serve_slots,
                       slice_indices=slice_indices, variant=args.serve_variant)
        mark_run()
        t_start = time.perf_counter()
        lat = srv.replay(arrivals, (ptrs if on_device else host_ptrs)[:k], sizes[:k], on_device=on_device,
                         release_on_done=True)
        wall = time.perf_counter() - t_start
        st = srv.stats()
        srv.close()
        return lat, wall, st

    run(np.zeros(min(R, 256)))  # warm-up
    lat, wall, st = run(np.zeros(R))
    sat_rate = R / wall
    out = {
        "workload": "table 2^24 x 32 B; requests log-uniform 2^8..2^18 indices, Zipf(1.1) rows; 4 GiB budget",
        "server": {"slots": args.serve_slots, "slice_indices": slice_indices, "variant": args.serve_variant},
        "requests": R, "indices": int(sizes.sum()),
        "saturation": {"req_per_s": round(sat_rate, 1), "gbs": round(total_bytes / wall / 1e9, 1),
                       "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "steps": st["steps"]},
    }
    # Slices of 2^16 indices: requests above 2^16 span several steps (the
    # continuous-batching path proper); saturation and latency beside the
    # one-step configuration above.
    lat, wall, st = run(np.zeros(R), slice_indices=1 << 16)
    out["saturation_slice_2p16"] = {
        "req_per_s": round(R / wall, 1), "gbs": round(total_bytes / wall / 1e9, 1),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "steps": st["steps"],
        "requests_spanning_steps": int((sizes > (1 << 16)).sum())}
    # Indices handed over in host memory: the rate includes the host-side
    # check and copy and the upload of 4 bytes per index over PCIe.
    lat, wall, st = run(np.zeros(R), on_device=False)
    host_rate = R / wall
    out["saturation_host_indices"] = {
        "req_per_s": round(host_rate, 1), "gbs": round(total_bytes / wall / 1e9, 1),
        "index_upload_gbs": round(4 * int(sizes.sum()) / wall / 1e9, 2),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "steps": st["steps"]}
    rng = np.random.default_rng(args.seed + 13)
    arrivals = np.cumsum(rng.exponential(1.0 / (0.5 * host_rate), size=R))
    lat, wall, st = run(arrivals, on_device=False)
    out["poisson_host_indices_50pct"] = {
        "rate_req_per_s": round(0.5 * host_rate, 1), "gbs": round(total_bytes / wall / 1e9, 1),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "steps": st["steps"]}
    for frac in (0.25, 0.5, 0.8, 0.95):
        rate = frac * sat_rate
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=R))
        lat, wall, st = run(arrivals)
        out[f"poisson_{int(frac * 100)}pct"] = {
            "rate_req_per_s": round(rate, 1), "gbs": round(total_bytes / wall / 1e9, 1),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "steps": st["steps"]}
    # The same arrivals at fixed absolute rates, comparable across builds
    # whose saturation rates differ.
    for rate in (5e5, 1e6, 1.5e6):
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=R))
        lat, wall, st = run(arrivals)
        out[f"poisson_{rate / 1e6:g}M_per_s"] = {
            "rate_req_per_s": rate, "gbs": round(total_bytes / wall / 1e9, 1),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "steps": st["steps"]}
    # Bit-exactness: a few requests alone vs the oracle.
    import oracle

    mark_run()
    host_table = table.cpu().numpy().view(np.uint8)
    ok = True
    for i in range(0, R, max(1, R // 6))[:6]:
        srv = T.Server(table, n_rows, rb, budget_bytes=budget, max_slots=4, slice_indices=1 << 16)
        r = host_idx[offs[i]:offs[i] + sizes[i]]
        rid = srv.submit(r, r.size, on_device=False)
        srv.drain()
        ok = ok and np.array_equal(srv.fetch(rid, r.size), oracle.gather(host_table, r, rb))
        srv.close()
    out["bit_exact_sampled"] = bool(ok)
    return out


def _counter_rows(directory: str):
    """Rows of every rocprofv3 counter_collection.csv under directory."""
    import csv
    import glob

    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no *counter_collection.csv under {directory}")
    return rows


def _per_leg(rows, counter: str, n_legs: int):
    """Counter value summed over the library's dispatches of each leg: a
    marker (k_memcpy16 of grid size 256) opens the next leg; dispatches
    before the first marker (input set-up) and outside the library (torch's
    own kernels) are not counted."""
    disp = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        rec = disp.setdefault(d, [r["Kernel_Name"], int(r["Grid_Size"]), 0.0])
        rec[2] += float(r["Counter_Value"])
    out = [0.0] * n_legs
    leg = -1
    for d in sorted(disp):
        name, grid, value = disp[d]
        if "k_memcpy16" in name and grid == 256:
            leg += 1
            continue
        if leg >= 0 and "trove_amd::" in name:
            if leg >= n_legs:
                raise SystemExit("more marked legs in the counter run than in the leg list")
            out[leg] += value
    if leg + 1 != n_legs:
        raise SystemExit(f"{leg + 1} marked legs in the counter run, {n_legs} in the leg list")
    return out



def _trace_rows(directory: str):
    """(kernel name, grid threads, start ns, end ns) of every dispatch in the
    rocprofv3 kernel_trace.csv files under directory, in start order."""
    import csv
    import glob

    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", "*kernel_trace.csv"), recursive=True)):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                grid = 1
                for axis in ("X", "Y", "Z"):
                    grid *= int(r.get(f"Grid_Size_{axis}") or 1)
                if "Grid_Size" in r and r["Grid_Size"]:
                    grid = int(r["Grid_Size"])
                rows.append((r["Kernel_Name"], grid, int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    if not rows:
        raise SystemExit(f"no *kernel_trace.csv under {directory}")
    rows.sort(key=lambda x: x[2])
    return rows


def _short_kernel(name: str) -> str:
    """A kernel's demangled name without its return type, namespace and
    parameter list."""
    name = name.split("(")[0] if "(" in name else name
    for prefix in ("void ",):
        if name.startswith(prefix):
            name = name[len(prefix):]
    return name.replace("trove_amd::", "").replace("trove::", "").replace("unsigned int", "u32").strip()


# Sweep kernels and their algorithmic bytes per struct of S bytes.
_SWEEP_KERNELS = {"k_copy_shfl": lambda s: 2 * s, "k_gather_tile": lambda s: 2 * s + 4,
                  "k_gather_shfl": lambda s: 2 * s + 4}


def trace_from(directory: str, timeline=None, n_log2: int = 26) -> int:
    """Markdown summary of a kernel trace: per (kernel, grid size) launches
    and durations, then the struct-size sweep's kernels (their dispatches at
    the largest grid of each, the sweep's launches) as algorithmic GB/s at
    N = 2^n_log2; or, with a timeline prefix, those dispatches in time order
    with the gap before each (runs split at pauses over 2 ms)."""
    import re
    import statistics as st

    rows = _trace_rows(directory)
    if timeline:
        # Runs are split at marker dispatches (a one-block k_memcpy16, grid
        # 256) where the program placed them, else at pauses over 2 ms.
        def is_marker(r):
            return "k_memcpy16" in r[0] and r[1] == 256

        markers = any(is_marker(r) for r in rows)
        sel = [r for r in rows if _short_kernel(r[0]).startswith(timeline) or (markers and is_marker(r))]
        if not any(not is_marker(r) for r in sel):
            raise SystemExit(f"no dispatch of a kernel starting with {timeline!r}")
        run, prev_end, started = 0, None, False
        print("| run | step | kernel | grid (threads) | duration us | gap before us |")
        print("|---|---|---|---|---|---|")
        step = 0
        for r in sel:
            name, grid, t0, t1 = r
            if is_marker(r):
                if started:
                    run, step, prev_end = run + 1, 0, None
                continue
            started = True
            gap = None if prev_end is None else (t0 - prev_end) / 1e3
            if not markers and gap is not None and gap > 2000:
                run, step, gap = run + 1, 0, None
            print(f"| {run} | {step} | `{_short_kernel(name).split('<')[0]}` | {grid} | {(t1 - t0) / 1e3:.1f} | "
                  f"{'' if gap is None else f'{gap:.1f}'} |")
            prev_end, step = t1, step + 1
        return 0
    groups = {}
    for name, grid, t0, t1 in rows:
        groups.setdefault((_short_kernel(name), grid), []).append((t1 - t0) / 1e3)
    total_us = sum(sum(v) for v in groups.values())
    print("| kernel | grid (threads) | launches | avg us | min us | max us | total ms | share |")
    print("|---|---|---|---|---|---|---|---|")
    for (name, grid), d in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{name}` | {grid} | {len(d)} | {st.mean(d):.1f} | {min(d):.1f} | {max(d):.1f} | "
              f"{sum(d) / 1e3:.2f} | {100 * sum(d) / total_us:.1f}% |")
    sweep = {}
    for (name, grid), d in groups.items():
        m = re.match(r"(k_\w+)<array<u32, (\d+)>", name)
        # The sweep's sizes are k = 1..16 words; the large-struct legs share
        # the kernel names but run at other N.
        if not m or m.group(1) not in _SWEEP_KERNELS or int(m.group(2)) > 16:
            continue
        key = (m.group(1), int(m.group(2)))
        if key not in sweep or grid > sweep[key][0]:
            sweep[key] = (grid, d)
    if sweep:
        n = 1 << n_log2
        print()
        print(f"Sweep kernels (their dispatches at the largest grid of each), N = 2^{n_log2} structs:")
        print()
        print("| kernel | k | grid (threads) | launches | avg us | algorithmic GB/s |")
        print("|---|---|---|---|---|---|")
        for (kern, k), (grid, d) in sorted(sweep.items(), key=lambda kv: (kv[0][1], kv[0][0])):
            gbs = _SWEEP_KERNELS[kern](4 * k) * n / (st.mean(d) * 1e-6) / 1e9
            print(f"| `{kern}` | {k} | {grid} | {len(d)} | {st.mean(d):.1f} | {gbs:.0f} |")
    return 0

def counters_from(dirs, legs_json: str) -> int:
    """Per-leg means of every counter found in the counter runs under dirs
    (each a --pmc run of the same --profile-legs program), as markdown."""
    with open(legs_json) as f:
        legs = json.load(f)["legs"]
    table = {}  # counter -> per-leg values
    for d in dirs:
        rows = _counter_rows(d)
        for counter in sorted({r["Counter_Name"] for r in rows}):
            table[counter] = _per_leg(rows, counter, len(legs))
    counters = sorted(table)
    per = {}
    for i, (name, _) in enumerate(legs):
        rec = per.setdefault(name, [0] + [0.0] * len(counters))
        rec[0] += 1
        for c, counter in enumerate(counters):
            rec[1 + c] += table[counter][i]
    print("| leg | " + " | ".join(counters) + " |")
    print("|---" * (len(counters) + 1) + "|")
    for name in sorted(per):
        rec = per[name]
        print(f"| `{name}` | " + " | ".join(f"{v / rec[0]:.4g}" for v in rec[1:]) + " |")
    return 0


def traffic_from(fetch_dir: str, write_dir: str, legs_json: str, out_path: str = TRAFFIC_PATH) -> int:
    """Builds profiles/traffic.json (HBM bytes per launch of every bench leg)
    from FETCH_SIZE and WRITE_SIZE counter runs of `bench.py --profile-legs`.
    Calibration: the memcpy_ceiling leg copies a known number of bytes each
    way; FETCH_SIZE and WRITE_SIZE are scaled by what they read for it."""
    import subprocess

    with open(legs_json) as f:
        legs = json.load(f)["legs"]
    fetch = _per_leg(_counter_rows(fetch_dir), "FETCH_SIZE", len(legs))
    write = _per_leg(_counter_rows(write_dir), "WRITE_SIZE", len(legs))
    per = {}
    for (name, nbytes), fk, wk in zip(legs, fetch, write):
        rec = per.setdefault(name, [0, 0.0, 0.0, nbytes])
        rec[0] += 1
        rec[1] += fk
        rec[2] += wk
    if "memcpy_ceiling" not in per:
        raise SystemExit("the leg list has no memcpy_ceiling leg to calibrate on")
    c = per["memcpy_ceiling"]
    each_way = c[3] / 2
    f_fac = each_way / (c[1] / c[0] * 1024)
    w_fac = each_way / (c[2] / c[0] * 1024)
    try:
        paths = [f"trove_amd/csrc/{f}" for f in TRAFFIC_SOURCES] + ["include/trove"]
        commit = subprocess.run(["git", "-C", ROOT, "log", "-1", "--format=%h", "--", *paths],
                                capture_output=True, text=True, timeout=20).stdout.strip() or None
    except (OSError, subprocess.SubprocessError):
        commit = None
    kernels, reads, writes, table = {}, {}, {}, []
    for name in sorted(per):
        n_, fk, wk, nbytes = per[name]
        fk, wk = fk / n_, wk / n_
        rd, wr = fk * 1024 * f_fac, wk * 1024 * w_fac
        kernels[name] = int(round(rd + wr))
        reads[name], writes[name] = int(round(rd)), int(round(wr))
        table.append(f"| `{name}` | {fk:.0f} | {wk:.0f} | {rd / 1e9:.3f} | {wr / 1e9:.3f} | {nbytes / 1e9:.3f} | "
                     f"{(rd + wr) / nbytes:.2f} |")
    out = {"unit": "HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated on the plain copy)",
           "kernel_sources_sha256": kernel_sources_sha256(), "kernels_commit": commit,
           "calibration": {"fetch_factor": round(f_fac, 4), "write_factor": round(w_fac, 4),
                           "bytes_each_way": int(each_way)},
           "kernels": kernels, "read": reads, "written": writes}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"Kernel sources sha256 {out['kernel_sources_sha256'][:16]} (commit {commit}); calibration on the "
          f"memcpy_ceiling leg ({int(each_way)} bytes each way): FETCH_SIZE x {f_fac:.3f}, WRITE_SIZE x {w_fac:.3f}.")
    print()
    print("| leg | FETCH KiB | WRITE KiB | read GB (cal.) | written GB (cal.) | algorithmic GB | HBM / algorithmic |")
    print("|---|---|---|---|---|---|---|")
    print("\n".join(table))
    return 0


def load_traffic(path: str = TRAFFIC_PATH, split: bool = False):
    """(kernels, stale): the committed per-leg HBM bytes, and whether they were
    measured on other kernel sources than these (then kernels is None). With
    split, kernels maps each leg to (read bytes, written bytes)."""
    if not os.path.exists(path):
        return None, True
    with open(path) as f:
        tr = json.load(f)
    if tr.get("kernel_sources_sha256") != kernel_sources_sha256():
        return None, True
    if split:
        rd, wr = tr.get("read", {}), tr.get("written", {})
        return {k: (rd[k], wr[k]) for k in rd if k in wr}, False
    return tr.get("kernels", {}), False


def hbm_bound_ms(read_bytes, written_bytes, index_bytes, random_rows, random_gbs, stream_gbs):
    """Time the HBM needs for a launch's measured traffic: the random row
    fills (the read bytes beyond the index stream, when random_rows) at the
    random-line rate, everything else (index stream, contiguous reads, the
    writes) at the streaming-copy rate."""
    random = max(read_bytes - index_bytes, 0) if random_rows else 0
    streaming = read_bytes + written_bytes - random
    return (random / (random_gbs * 1e9) + streaming / (stream_gbs * 1e9)) * 1e3



def bench_line_order(args, torch, T):
    """2^26 random 128-byte line reads from a 4 GiB source (the random-line
    ceiling kernel), with the same line indices in random order and sorted
    within consecutive windows of 64 to 2^20 indices, then fully sorted:
    whether an order a gather could stage (DRAM page locality, address
    translation) raises the rate random row fills are served at."""
    import numpy as np

    dev = torch.device("cuda:0")
    n_lines = (4 << 30) // 128
    n_reads = 1 << 26
    src = torch.empty(n_lines * 32, dtype=torch.int32, device=dev)
    T.fill_words_device(src, src.numel(), args.seed + 40)
    out = torch.empty(n_reads, dtype=torch.int32, device=dev)
    base = T.gen_indices(n_reads, n_lines, "random", args.seed + 41)
    rates = {}
    for w in ("random", 64, 1 << 10, 1 << 14, 1 << 18, 1 << 20, "sorted"):
        if w == "random":
            order = base
        elif w == "sorted":
            order = np.sort(base)
        else:
            order = np.sort(base.reshape(-1, w), axis=1).reshape(-1)
        d_idx = torch.from_numpy(np.ascontiguousarray(order)).to(dev)
        T.random_lines_ceiling(src, n_lines, d_idx, out, n_reads)
        ms = statistics.median(timed_runs(torch, lambda: T.random_lines_ceiling(src, n_lines, d_idx, out, n_reads), 5))
        rates[str(w)] = round(128 * n_reads / (ms * 1e-3) / 1e9, 1)
        log(f"bench line-order: window {w}: {rates[str(w)]} GB/s")
        del d_idx
    return {"workload": "2^26 reads of random 128-byte lines of a 4 GiB source; indices in random order, "
                        "sorted within windows of the given length, or fully sorted; median of 5",
            "gbs_by_window": rates}

def main():
    args = parse_args()
    if args.traffic_from:
        return traffic_from(*args.traffic_from)
    if args.trace_from:
        return trace_from(args.trace_from, args.timeline, args.n_log2)
    if args.counters_from:
        if len(args.counters_from) < 2:
            raise SystemExit("--counters-from needs at least one counter directory and the legs JSON")
        return counters_from(args.counters_from[:-1], args.counters_from[-1])
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        log("bench.py: only --gpus 1 is supported (one MI355X); see the sharding design notes")
        return 2

    import torch

    import trove_amd as T

    T.lib()  # fail loudly if the HIP library is missing
    assert torch.cuda.is_available(), "no GPU visible"
    profiling = args.profile_legs is not None
    if profiling:
        ProfileLegs.legs = []
    if args.config == "batched":
        print(json.dumps({"metric": "batched gather jobs/s", "batched": bench_batched(args, torch, T)}), flush=True)
        return 0
    if args.config == "served":
        print(json.dumps({"metric": "served gather p99 latency", "served": bench_served(args, torch, T)}),
              flush=True)
        return 0
    if args.config == "line-order":
        print(json.dumps({"metric": "random 128-byte line reads GB/s", "line_order": bench_line_order(args, torch, T)}),
              flush=True)
        return 0
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

    clock = KernelClock(torch, seen=warm.seen)
    if args.warmup == 0:  # load the code objects outside the timed region
        step(KernelClock(torch, seen=clock.seen))
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(clock)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    clock.collect()
    log(f"bench: timed {args.steps} steps in {wall:.3f}s")

    per = {}
    if profiling:  # no timings: the legs run once each, marked
        dname = None
        value = ms_per_step = 0.0
    else:
        step_bytes = sum(v[2] for v in clock.totals.values()) / args.steps
        ms_per_step = wall * 1e3 / args.steps
        value = step_bytes / (ms_per_step * 1e-3) / 1e9
    for name, (launches, ms, nbytes) in clock.totals.items():
        avg_ms = ms / launches
        med_ms = statistics.median(clock.samples[name])
        per[name] = {"avg_ms": round(avg_ms, 4), "median_ms": round(med_ms, 4),
                     "gbs": round(nbytes / launches / (avg_ms * 1e-3) / 1e9, 1),
                     "median_gbs": round(nbytes / launches / (med_ms * 1e-3) / 1e9, 1)}
    if not profiling:
        dominant = max(clock.totals.items(), key=lambda kv: kv[1][1])
        dname, (dl, dms, dbytes) = dominant
        d_gbs = dbytes / dl / (dms / dl * 1e-3) / 1e9
    r3, r2 = (1, 1) if profiling else (3, 2)  # launches per leg

    # Naive per-lane kernels and the plain-copy ceiling, same inputs, untimed
    # by the step clock.
    naive = {}
    if not args.no_naive:
        nclock = KernelClock(torch)
        sums = torch.empty(n, dtype=torch.int32, device=dev)  # the load benchmark's per-struct sums
        ws = torch.empty(n, dtype=torch.int32, device=dev)  # trove_scatter_permutation's workspace
        for k in ks:
            s = 4 * k
            for _ in range(r3):
                nclock.run(f"copy_direct_k{k}", 2 * s * n, lambda: T.copy(src, dst, n, s, variant="direct"))
                nclock.run(f"gather_direct_k{k}", (2 * s + 4) * n,
                           lambda: T.gather(src, idx, dst, n, s, variant="direct"))
                # The naive per-lane access launched as the library's own
                # kernels are: what the transposition adds on equal terms.
                nclock.run(f"copy_direct_launch_k{k}", 2 * s * n,
                           lambda: T.copy(src, dst, n, s, variant="direct_launch"))
                nclock.run(f"gather_direct_launch_k{k}", (2 * s + 4) * n,
                           lambda: T.gather(src, idx, dst, n, s, variant="direct_launch"))
                # The LDS-staged variant (trove/lds.h) beside the transposition.
                nclock.run(f"copy_lds_k{k}", 2 * s * n, lambda: T.copy(src, dst, n, s, variant="lds"))
                nclock.run(f"gather_lds_k{k}", (2 * s + 4) * n, lambda: T.gather(src, idx, dst, n, s, variant="lds"))
            for _ in range(r2):
                nclock.run(f"scatter_tile_k{k}", (2 * s + 4) * n, lambda: T.scatter(src, idx, dst, n, s, variant="tile"))
                nclock.run(f"scatter_shfl_k{k}", (2 * s + 4) * n, lambda: T.scatter(src, idx, dst, n, s, variant="shfl"))
                nclock.run(f"scatter_direct_k{k}", (2 * s + 4) * n,
                           lambda: T.scatter(src, idx, dst, n, s, variant="direct"))
                nclock.run(f"scatter_direct_launch_k{k}", (2 * s + 4) * n,
                           lambda: T.scatter(src, idx, dst, n, s, variant="direct_launch"))
                nclock.run(f"scatter_lds_k{k}", (2 * s + 4) * n, lambda: T.scatter(src, idx, dst, n, s, variant="lds"))
                # The permutation scatter (inversion into ws + gather where
                # that is faster; trove_scatter_permutation).
                nclock.run(f"scatter_perm_k{k}", (2 * s + 4) * n,
                           lambda: T.scatter_permutation(src, idx, dst, n, s, workspace=ws))
                # The reference benchmark's store-only and load-only kernels
                # (tests/benchmark.cu:52-88; bytes S and S + 4 per struct).
                for v in ("shfl", "direct"):
                    nclock.run(f"store_{v}_k{k}", s * n, lambda v=v: T.store_counting(dst, n, s, variant=v))
                    nclock.run(f"load_{v}_k{k}", (s + 4) * n, lambda v=v: T.load_sum(src, sums, n, s, variant=v))
        for _ in range(r3):
            nclock.run("memcpy_ceiling", 2 * n * smax, lambda: T.memcpy_ceiling(src, dst, n * smax))
        # Block interface: 2^28 32-bit elements, s per lane, converged and
        # with one wavefront group in eight diverged.
        e = 1 << 28
        block_fits = e <= src.numel() and e <= dst.numel()  # 1 GiB each way; the default sweep has 4 GiB
        if not block_fits:
            log(f"bench: block interface leg skipped (buffers sized for --k {args.k})")
        for sb in (1, 3, 5, 7, 9, 11, 13, 15) if block_fits else ():
            for _ in range(r2):
                nclock.run(f"block_s{sb}", 8 * e, lambda: T.block_copy(src, dst, e, sb))
                nclock.run(f"block_div_s{sb}", 8 * e, lambda: T.block_copy(src, dst, e, sb, divergent=True, seed=7))
        # Random row writes (rows of 32, 60, 64 and 128 bytes at the rows of a
        # permutation, 2^26 rows up to 64 bytes, 2^25 of 128 bytes; ordinary
        # and nontemporal stores): the rate the scatters' row writes are
        # bounded by.
        for rb in (32, 60, 64, 128):
            n_w = n if rb <= 64 else n // 2
            if n_w * rb > dst.numel() * 4 or n_w < 64:
                continue
            w_idx = idx if n_w == n else torch.from_numpy(T.gen_permutation(n_w, args.seed + 43)).to(dev)
            for _ in range(r3):
                for nt in (False, True):
                    nclock.run(f"write_ceiling_{rb}_{'nt' if nt else 'cached'}", rb * n_w,
                               lambda: T.random_rows_write_ceiling(dst, n_w, rb, w_idx, n_w, nontemporal=nt))
            del w_idx
        # Random whole 128-byte lines of the 4 GiB source: the rate the
        # random gathers' line fills are bounded by (last: it leaves the
        # source's pages scattered through the caches and TLBs).
        n_lines = n * smax // 128
        if n_lines >= 1024:
            n_reads = 1 << 26
            line_idx = torch.from_numpy(T.gen_indices(n_reads, n_lines, "random", args.seed + 41)).to(dev)
            line_out = torch.empty(n_reads, dtype=torch.int32, device=dev)
            for _ in range(r3):
                nclock.run("random_lines_ceiling", 128 * n_reads,
                           lambda: T.random_lines_ceiling(src, n_lines, line_idx, line_out, n_reads))
            del line_idx, line_out
        torch.cuda.synchronize()
        nclock.collect()
        del sums, ws
        for name, (launches, ms, nbytes) in nclock.totals.items():
            avg_ms = ms / launches
            naive[name] = {"avg_ms": round(avg_ms, 4), "gbs": round(nbytes / launches / (avg_ms * 1e-3) / 1e9, 1)}

    sweep = []
    for k in ks if not profiling else ():
        row = {"k": k, "bytes": 4 * k,
               "copy_gbs": per[f"copy_shfl_k{k}"]["gbs"],
               "gather_tile_gbs": per[f"gather_tile_k{k}"]["gbs"],
               "gather_coalesced_ptr_gbs": per[f"gather_shfl_k{k}"]["gbs"],
               "median_gbs": {"copy": per[f"copy_shfl_k{k}"]["median_gbs"],
                              "gather_tile": per[f"gather_tile_k{k}"]["median_gbs"],
                              "gather_coalesced_ptr": per[f"gather_shfl_k{k}"]["median_gbs"]}}
        if naive:
            cd, gd = naive[f"copy_direct_k{k}"]["gbs"], naive[f"gather_direct_k{k}"]["gbs"]
            row.update({"copy_naive_gbs": cd, "gather_naive_gbs": gd,
                        "copy_vs_naive": round(row["copy_gbs"] / cd, 2),
                        "gather_tile_vs_naive": round(row["gather_tile_gbs"] / gd, 2),
                        "gather_coalesced_ptr_vs_naive": round(row["gather_coalesced_ptr_gbs"] / gd, 2)})
            row.update({"scatter_tile_gbs": naive[f"scatter_tile_k{k}"]["gbs"],
                        "scatter_coalesced_ptr_gbs": naive[f"scatter_shfl_k{k}"]["gbs"],
                        "scatter_naive_gbs": naive[f"scatter_direct_k{k}"]["gbs"]})
            # The naive access in the library's launch (DIRECT_LAUNCH).
            cl, gl = naive[f"copy_direct_launch_k{k}"]["gbs"], naive[f"gather_direct_launch_k{k}"]["gbs"]
            row["naive_launch_equal"] = {
                "copy_gbs": cl, "gather_gbs": gl, "scatter_gbs": naive[f"scatter_direct_launch_k{k}"]["gbs"],
                "copy_vs": round(row["copy_gbs"] / cl, 2),
                "gather_tile_vs": round(row["gather_tile_gbs"] / gl, 2),
                "scatter_tile_vs": round(row["scatter_tile_gbs"] / naive[f"scatter_direct_launch_k{k}"]["gbs"], 2)}
            best_scatter = max(row["scatter_tile_gbs"], row["scatter_coalesced_ptr_gbs"], row["scatter_naive_gbs"],
                               naive[f"scatter_lds_k{k}"]["gbs"])
            row["scatter_permutation"] = {"gbs": naive[f"scatter_perm_k{k}"]["gbs"],
                                          "path": "invert + gather" if T.scatter_permutation_inverts(4 * k)
                                                  else "scatter",
                                          "vs_best_scatter": round(naive[f"scatter_perm_k{k}"]["gbs"] / best_scatter, 3)}
            row["lds_gbs"] = {"copy": naive[f"copy_lds_k{k}"]["gbs"], "gather": naive[f"gather_lds_k{k}"]["gbs"],
                              "scatter": naive[f"scatter_lds_k{k}"]["gbs"]}
            row["store_gbs"] = {"shfl": naive[f"store_shfl_k{k}"]["gbs"], "naive": naive[f"store_direct_k{k}"]["gbs"]}
            row["load_gbs"] = {"shfl": naive[f"load_shfl_k{k}"]["gbs"], "naive": naive[f"load_direct_k{k}"]["gbs"]}
            # The path AUTO takes at this size (LDS-staged or register
            # transposition) and its rate from that path's own leg above.
            av = T.auto_variant("copy", 4 * k)
            row["copy_auto"] = {"variant": av,
                                "gbs": row["copy_gbs"] if av == "shfl" else row["lds_gbs"]["copy"]}
            row["copy_auto"]["frac_peak"] = round(row["copy_auto"]["gbs"] / HBM_PEAK_GBS, 3)
        row["copy_frac_peak"] = round(row["copy_gbs"] / HBM_PEAK_GBS, 3)
        sweep.append(row)
    # Structs beyond the sweep: 96..1024 bytes (6..64 16-byte words; 96, 160,
    # 192, 224, 384 and 768 bytes take the composite schedule), N = 2^24 up
    # to 256 bytes and 2^22 above.
    large = None
    if not args.no_naive:
        lclock = KernelClock(torch)
        for nl, sizes in ((1 << 24, (96, 128, 160, 192, 224, 256)), (1 << 22, (384, 512, 768, 1024))):
            # Up to 4 GiB of rows = the size of src / dst at the default sweep
            # (k up to 16); indices are their own permutation of [0, nl).
            if nl * max(sizes) > src.numel() * 4 or nl * max(sizes) > dst.numel() * 4:
                log(f"bench: large structs {sizes[0]}..{sizes[-1]} B skipped (buffers sized for --k {args.k})")
                continue
            idx_l = torch.from_numpy(T.gen_permutation(nl, args.seed + 31)).to(dev)
            for sb in sizes:
                for _ in range(r3):
                    lclock.run(f"copy_{sb}", 2 * sb * nl, lambda: T.copy(src, dst, nl, sb, variant="shfl"))
                    lclock.run(f"copy_naive_{sb}", 2 * sb * nl, lambda: T.copy(src, dst, nl, sb, variant="direct"))
                    lclock.run(f"copy_naive_launch_{sb}", 2 * sb * nl,
                               lambda: T.copy(src, dst, nl, sb, variant="direct_launch"))
                    lclock.run(f"gather_{sb}", (2 * sb + 4) * nl,
                               lambda: T.gather(src, idx_l, dst, nl, sb, variant="tile"))
                    lclock.run(f"gather_coalesced_ptr_{sb}", (2 * sb + 4) * nl,
                               lambda: T.gather(src, idx_l, dst, nl, sb, variant="shfl"))
                    lclock.run(f"gather_naive_{sb}", (2 * sb + 4) * nl,
                               lambda: T.gather(src, idx_l, dst, nl, sb, variant="direct"))
                    lclock.run(f"gather_naive_launch_{sb}", (2 * sb + 4) * nl,
                               lambda: T.gather(src, idx_l, dst, nl, sb, variant="direct_launch"))
                    if T.lds_supported(sb):
                        lclock.run(f"copy_lds_{sb}", 2 * sb * nl, lambda: T.copy(src, dst, nl, sb, variant="lds"))
                        lclock.run(f"gather_lds_{sb}", (2 * sb + 4) * nl,
                                   lambda: T.gather(src, idx_l, dst, nl, sb, variant="lds"))
            torch.cuda.synchronize()
            del idx_l
        lclock.collect()
        large = {"workload": "struct sizes 96..1024 B (composite schedule at 96/160/192/224/384/768), "
                             "N=2^24 up to 256 B and 2^22 above, gather through a seeded permutation of [0, N)"}
        for name, (launches, ms, nbytes) in lclock.totals.items():
            large[name + "_gbs"] = round(nbytes / launches / (ms / launches * 1e-3) / 1e9, 1)
    if profiling:
        with open(args.profile_legs, "w") as f:
            json.dump({"legs": ProfileLegs.legs, "kernel_sources_sha256": kernel_sources_sha256()}, f)
        log(f"bench: {len(ProfileLegs.legs)} marked legs written to {args.profile_legs}")
        return 0
    block = None
    if naive and "block_s1" in naive:
        block = {"workload": "block interface copy, 2^28 x 32-bit elements, s per lane; divergent: 1 in 8 "
                             "64-array groups with lanes masked (warp_converged fallback)",
                 "gbs": {f"s{sb}": naive[f"block_s{sb}"]["gbs"] for sb in (1, 3, 5, 7, 9, 11, 13, 15)},
                 "divergent_gbs": {f"s{sb}": naive[f"block_div_s{sb}"]["gbs"] for sb in (1, 3, 5, 7, 9, 11, 13, 15)}}

    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (profiles/traffic.json; null when those
    # were measured on other kernel sources than these).
    traffic_table, traffic_stale = load_traffic(split=True)
    rw = traffic_table.get(dname) if traffic_table else None
    traffic = rw[0] + rw[1] if rw else None
    hbm_gbs = traffic / (dms / dl * 1e-3) / 1e9 if traffic else None
    # The measured ceilings matching the dominant kernel's traffic: a
    # gather's random row fills at the random-line read rate, its index
    # stream and output writes (a copy's whole traffic) at the streaming-copy
    # rate; ceiling_frac = that bound's time / the kernel's time.
    random_gbs = naive.get("random_lines_ceiling", {}).get("gbs")
    stream_gbs = naive.get("memcpy_ceiling", {}).get("gbs")
    is_gather = dname.startswith("gather")
    bound_ms = None
    if rw and stream_gbs and (random_gbs or not is_gather):
        bound_ms = hbm_bound_ms(rw[0], rw[1], 4 * n if is_gather else 0, is_gather, random_gbs or 1.0, stream_gbs)
    ceiling_gbs = round(traffic / (bound_ms * 1e-3) / 1e9, 1) if bound_ms else None
    # Random-row write ceilings and the scatters of those row sizes against
    # them (row bytes written per second by the tile scatter / by the ceiling
    # kernel's better store policy).
    write_ceiling = {}
    for rb in (32, 60, 64, 128):
        c, nt = naive.get(f"write_ceiling_{rb}_cached"), naive.get(f"write_ceiling_{rb}_nt")
        if c and nt:
            write_ceiling[str(rb)] = {"cached": c["gbs"], "nontemporal": nt["gbs"], "best": max(c["gbs"], nt["gbs"])}
    scatter_vs_ceiling = {}
    memcpy_gbs = naive.get("memcpy_ceiling", {}).get("gbs")
    for k in ks:
        rb = 4 * k
        leg = naive.get(f"scatter_tile_k{k}")
        if str(rb) in write_ceiling and leg:
            ceiling = write_ceiling[str(rb)]["best"]
            row_gbs = rb * n / (leg["avg_ms"] * 1e-3) / 1e9
            rec = {"kernel": f"scatter_tile_k{k}", "row_write_gbs": round(row_gbs, 1), "ceiling_gbs": ceiling,
                   "ceiling_frac": round(row_gbs / ceiling, 3)}
            if memcpy_gbs:
                # The scatter also streams its source rows and indices: bound =
                # the row writes at the write ceiling + S + 4 streamed bytes per
                # row at the copy rate.
                sc_ms = (rb * n / (ceiling * 1e9) + (rb + 4) * n / (memcpy_gbs * 1e9)) * 1e3
                rec.update(bound_ms=round(sc_ms, 3), bound_frac=round(sc_ms / leg["avg_ms"], 3))
            scatter_vs_ceiling[str(rb)] = rec

    # Contiguous copy, small: 20-byte structs, N = 2^20 (21 MB: resident in
    # the 256 MB Infinity Cache, so this is a warm-cache number), checked
    # bit-exact against the scalar host copy, which is also its CPU path.
    small = None
    if not args.no_cpu:
        import numpy as np

        import oracle

        ns, ss = 1 << 20, 20
        words = T.gen_words(ns * ss // 4, args.seed + 21)
        s_src = torch.from_numpy(words.view(np.int32)).to(dev)
        s_dst = torch.zeros_like(s_src)
        # Warm: 20 launches back to back in one event bracket (the device's
        # rate with the working set cached); single: one launch per bracket
        # after a synchronisation, which starts every 12 us copy from an idle
        # GPU.
        def back_to_back():
            for _ in range(20):
                T.copy(s_src, s_dst, ns, ss, variant="shfl")
        t_gpu = statistics.median(timed_runs(torch, back_to_back, 5)) / 20
        t_single = statistics.median(timed_runs(torch, lambda: T.copy(s_src, s_dst, ns, ss, variant="shfl"), 20))
        # Flushed: 1 GiB of writes between launches evicts the 42 MB working
        # set from the Infinity Cache; only the copy is timed.
        flush = torch.empty(1 << 28, dtype=torch.int32, device=dev)
        cold = []
        for r in range(20):
            T.fill_words_device(flush, flush.numel(), r)
            cold.extend(timed_runs(torch, lambda: T.copy(s_src, s_dst, ns, ss, variant="shfl"), 1))
        t_cold = statistics.median(cold)
        del flush
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        want = oracle.copy(words.view(np.uint8), ns, ss)
        t_cpu = time.perf_counter() - t0
        small = {"workload": "contiguous copy, 20-byte structs, N=2^20; warm (fits the Infinity Cache; 20 "
                             "launches back to back, median of 5), single (one launch per event bracket after "
                             "a synchronisation, median of 20) and flushed (1 GiB written before each of 20 "
                             "launches, median)",
                 "gpu_gbs": round(2 * ss * ns / (t_gpu * 1e-3) / 1e9, 1),
                 "gpu_single_launch_gbs": round(2 * ss * ns / (t_single * 1e-3) / 1e9, 1),
                 "gpu_flushed_gbs": round(2 * ss * ns / (t_cold * 1e-3) / 1e9, 1),
                 "cpu_scalar_gbs": round(2 * ss * ns / t_cpu / 1e9, 2),
                 "bit_exact": bool(np.array_equal(s_dst.cpu().numpy().view(np.uint8), want))}
        del s_src, s_dst
    del src, dst, idx
    torch.cuda.empty_cache()
    extra = {}
    if small:
        extra["small_contiguous"] = small
    if args.config == "all":
        log("bench: batched jobs")
        extra["batched"] = bench_batched(args, torch, T)
        torch.cuda.empty_cache()
        log("bench: served gather")
        extra["served"] = bench_served(args, torch, T)
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, OpenMP)")
        cpu = cpu_baseline(ks, args.cpu_n_log2, args.seed)

    out = {
        "metric": "AoS copy/gather GB/s (% HBM peak) vs naive loads; batched jobs/s, p99 latency",
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
            "traffic_stale": traffic_stale,
            "hbm_gbs": round(hbm_gbs, 1) if hbm_gbs else None,
            "ceiling": ("random row fills at random_line_ceiling_gbs + index and output streams at "
                        "memcpy_ceiling_gbs" if is_gather else "traffic at memcpy_ceiling_gbs"),
            "ceiling_ms": round(bound_ms, 4) if bound_ms else None,
            "ceiling_gbs": ceiling_gbs,
            "ceiling_frac": round(bound_ms / (dms / dl), 3) if bound_ms else None,
        },
        "cpu_baseline": cpu,
        "memcpy_ceiling_gbs": naive.get("memcpy_ceiling", {}).get("gbs"),
        "random_line_ceiling_gbs": naive.get("random_lines_ceiling", {}).get("gbs"),
        "random_write_ceiling_gbs": write_ceiling or None,
        "scatter_vs_write_ceiling": scatter_vs_ceiling or None,
        "block": block,
        "large_structs": large,
        "sweep": sweep,
        **extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
