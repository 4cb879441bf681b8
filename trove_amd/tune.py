"""Autotuner for batched gathers.

Picks, per (struct size, index pattern, length bucket), the kernel variant
(tile path, coalesced_ptr path with the shuffle transposition, its LDS-staged
counterpart, naive per-lane loads) and the work-item size
(64-struct groups per wavefront item) of the one-launch batch kernel, by
timing single-job batches of that shape on the device with HIP events. The
table is plain JSON so a tuned device can be reused without re-timing.
"""
from __future__ import annotations

import json
import math
import statistics

import trove_amd as T

SIZES = tuple(range(4, 65, 4))
PATTERNS = ("identity", "random", "runs")
LENGTHS = (1 << 10, 1 << 14, 1 << 18, 1 << 22)
CHUNKS = (1, 2, 4, 16, 64)


def bucket(n: int, lengths=LENGTHS) -> int:
    """The tuned length nearest to n on a log scale."""
    ln = math.log2(max(1, n))
    return min(lengths, key=lambda L: abs(math.log2(L) - ln))


def _variants(size: int):
    # One-word structs (4, 8, 16 B) have no transposition: the coalesced_ptr
    # path reduces to per-lane accesses (the tile path keeps its streaming
    # policies, the LDS-staged path turns the stores into dwordx4 chunks).
    return ("tile", "lds", "direct") if size in (4, 8, 16) else ("tile", "shfl", "lds", "direct")


def autotune(sizes=SIZES, patterns=PATTERNS, lengths=LENGTHS, reps: int = 3, seed: int = 1, log=None) -> dict:
    """Times every (variant, chunk) for every (size, pattern, length) and
    returns {"entries": [...]} with the fastest choice of each."""
    import torch

    dev = torch.device("cuda:0")
    nmax, smax = max(lengths), max(sizes)
    src = torch.empty(nmax * smax // 4, dtype=torch.int32, device=dev)
    dst = torch.empty_like(src)
    T.fill_words_device(src, src.numel(), seed)
    entries = []
    for pattern in patterns:
        for L in lengths:
            idx = torch.from_numpy(T.gen_indices(L, L, pattern, seed + L)).to(dev)
            for s in sizes:
                best = None
                for variant in _variants(s):
                    for chunk in (CHUNKS if variant != "shfl" else (1,)):  # shfl items are one group
                        b = T.Batch([(src, idx, dst, L, s, variant, chunk)])
                        b.run("launch")  # warm-up
                        times = []
                        for _ in range(reps):
                            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                            e0.record()
                            b.run("launch")
                            e1.record()
                            e1.synchronize()
                            times.append(e0.elapsed_time(e1))
                        b.close()
                        t = statistics.median(times)
                        if best is None or t < best[2]:
                            best = (variant, chunk, t)
                entries.append({"bytes": s, "pattern": pattern, "length": L, "variant": best[0],
                                "chunk_groups": best[1], "ms": round(best[2], 5)})
                if log:
                    log(f"tune {pattern:8s} L=2^{int(math.log2(L)):2d} S={s:2d}: {best[0]:6s} chunk={best[1]:2d} "
                        f"{best[2] * 1e3:8.1f} us")
    return {"version": 1, "lengths": list(lengths), "entries": entries}


def save(table: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def load(path: str) -> dict:
    with open(path) as f:
        return json.load(f)


def choose(table: dict, size: int, pattern: str, n: int):
    """(variant, chunk_groups) the table picks for a job; the tile path with
    16-group items when the table has no entry for the size or pattern."""
    lengths = tuple(table.get("lengths", LENGTHS))
    L = bucket(n, lengths)
    for e in table["entries"]:
        if e["bytes"] == size and e["pattern"] == pattern and e["length"] == L:
            return e["variant"], e["chunk_groups"]
    return "tile", 16
