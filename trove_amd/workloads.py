"""The batched-jobs and served-gather workloads, built the same way for the
benchmark and for the full-shape parity tests.

Batched: n_jobs independent gather jobs, struct size uniform over
{4, 8, ..., 64} bytes, length log-uniform on [2^10, 2^22] structs, index
pattern uniform over identity / uniform random permutation / sorted runs of
geometric length (mean 32); each job's source has as many structs as the job
(so the random pattern is a permutation of it).

Served: a device-resident table of 2^24 rows of 32 bytes (8 x uint32) and
requests of log-uniform 2^8..2^18 indices drawn Zipf(1.1) over the table.

Inputs are seeded and deterministic; payloads come from the library's device
fill (the same words trove_amd.gen_words gives on the host).
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import trove_amd as T

BATCH_MIN_LEN = 1 << 10
BATCH_MAX_LEN = 1 << 22
PATTERNS = ("identity", "random", "runs")

SERVE_ROWS = 1 << 24
SERVE_ROW_BYTES = 32
SERVE_MIN_REQ = 1 << 8
SERVE_MAX_REQ = 1 << 18
SERVE_ZIPF_S = 1.1


def _align(x: int, a: int = 256) -> int:
    return (x + a - 1) // a * a


def _threaded(fn, items, workers: int = 16):
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, items))


class BatchedWorkload:
    """The jobs of the batched config, resident in device memory: one source
    buffer holding every job's structs (each job 256-byte aligned), one index
    buffer, and two destination buffers of the source's layout."""

    def __init__(self, n_jobs: int, seed: int, torch, device="cuda:0"):
        rng = np.random.default_rng(seed + 3)
        self.n_jobs = n_jobs
        self.sizes = (4 * rng.integers(1, 17, size=n_jobs)).astype(np.int64)
        self.lengths = T.gen_loguniform(n_jobs, BATCH_MIN_LEN, BATCH_MAX_LEN, seed + 4)
        self.patterns = [PATTERNS[p] for p in rng.integers(0, 3, size=n_jobs)]
        self.src_off, self.idx_off = [], []
        tot_src = tot_idx = 0
        for s, L in zip(self.sizes, self.lengths):
            self.src_off.append(tot_src)
            tot_src += _align(int(s) * int(L))
            self.idx_off.append(tot_idx)
            tot_idx += _align(4 * int(L)) // 4
        t0 = time.time()
        self.src = torch.empty(tot_src // 4, dtype=torch.int32, device=device)
        self.dst_a = torch.empty_like(self.src)
        self.dst_b = torch.empty_like(self.src)
        T.fill_words_device(self.src, self.src.numel(), seed + 5)
        self.host_idx = np.empty(tot_idx, dtype=np.int32)

        def gen(j):
            L = int(self.lengths[j])
            o = self.idx_off[j]
            self.host_idx[o:o + L] = T.gen_indices(L, L, self.patterns[j], seed * 7919 + j)

        _threaded(gen, range(n_jobs))
        self.idx = torch.from_numpy(self.host_idx).to(device)
        self.setup_s = time.time() - t0

    @property
    def structs(self) -> int:
        return int(self.lengths.sum())

    def algorithmic_bytes(self) -> int:
        return int(sum((2 * int(s) + 4) * int(L) for s, L in zip(self.sizes, self.lengths)))

    def jobs(self, choose, dst):
        """Job tuples for trove_amd.Batch into `dst` (dst_a or dst_b);
        choose(size, pattern, length) -> (variant, chunk_groups)."""
        base, ibase, dbase = self.src.data_ptr(), self.idx.data_ptr(), dst.data_ptr()
        out = []
        for j in range(self.n_jobs):
            s, L = int(self.sizes[j]), int(self.lengths[j])
            v, c = choose(s, self.patterns[j], L)
            out.append((base + self.src_off[j], ibase + 4 * self.idx_off[j], dbase + self.src_off[j], L, s, v, c))
        return out

    def job_words(self, j: int):
        """(first word, word count) of job j's structs in src / dst."""
        return self.src_off[j] // 4, int(self.sizes[j]) * int(self.lengths[j]) // 4

    def job_indices(self, j: int) -> np.ndarray:
        o = self.idx_off[j]
        return self.host_idx[o:o + int(self.lengths[j])]


class ServedWorkload:
    """The served config's table (device) and a request stream: request i
    gathers host_idx[offs[i]:offs[i] + sizes[i]] (also resident on the device
    as idx)."""

    def __init__(self, n_requests: int, seed: int, torch, device="cuda:0", n_rows: int = SERVE_ROWS,
                 row_bytes: int = SERVE_ROW_BYTES):
        self.n_rows, self.row_bytes = n_rows, row_bytes
        t0 = time.time()
        self.table = torch.empty(n_rows * row_bytes // 4, dtype=torch.int32, device=device)
        T.fill_words_device(self.table, self.table.numel(), seed + 11)
        self.sizes = T.gen_loguniform(n_requests, SERVE_MIN_REQ, SERVE_MAX_REQ, seed + 12)
        self.offs = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        self.host_idx = np.empty(int(self.sizes.sum()), dtype=np.int32)

        def gen(i):
            o, n = int(self.offs[i]), int(self.sizes[i])
            self.host_idx[o:o + n] = T.gen_zipf(n, n_rows, SERVE_ZIPF_S, seed * 31 + i)

        _threaded(gen, range(n_requests))
        self.idx = torch.from_numpy(self.host_idx).to(device)
        self.setup_s = time.time() - t0

    @property
    def n_requests(self) -> int:
        return len(self.sizes)

    def request(self, i: int) -> np.ndarray:
        o = int(self.offs[i])
        return self.host_idx[o:o + int(self.sizes[i])]

    def device_request(self, i: int):
        o = int(self.offs[i])
        return self.idx[o:o + int(self.sizes[i])]

    def device_pointers(self) -> np.ndarray:
        return self.idx.data_ptr() + 4 * self.offs
