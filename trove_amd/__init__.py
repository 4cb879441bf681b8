"""Host-side interface to libtrove_amd, the MI355X AoS load/store library.

The device-side drop-in for Trove is the header library under include/trove/;
this module drives the C ABI of include/trove_amd.h (the launches the
reference's tests/benchmark.cu and tests/block.cu perform) through ctypes.
Buffers may be torch tensors on the GPU or raw device addresses (int). When
no stream is given, work is enqueued on torch's current stream, so torch
events and synchronisation see it.

The library is required: if libtrove_amd.so has not been built, importing
works but every call raises TroveUnavailable. There is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libtrove_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "trove_amd.h")

OK = 0
ERR_INVALID = -1
ERR_SIZE = -2
ERR_ALIGN = -3
ERR_VARIANT = -4
ERR_HIP = -5
ERR_NOMEM = -6
ERR_BUDGET = -7
ERR_FULL = -8
ERR_NOTFOUND = -9
ERR_AGAIN = -10
ERR_RANGE = -11

VARIANT_AUTO = 0
VARIANT_SHFL = 1
VARIANT_DIRECT = 2
VARIANT_TILE = 3

VARIANT_LDS = 4
VARIANT_DIRECT_LAUNCH = 5
VARIANTS = {"auto": VARIANT_AUTO, "shfl": VARIANT_SHFL, "direct": VARIANT_DIRECT, "tile": VARIANT_TILE,
            "lds": VARIANT_LDS, "direct_launch": VARIANT_DIRECT_LAUNCH}
# The LDS-staged variant takes struct sizes that are multiples of 4 bytes up
# to this many bytes (its stage is 64 structs per wavefront).
LDS_MAX_STRUCT_BYTES = 512


def lds_supported(struct_bytes: int) -> bool:
    return struct_bytes % 4 == 0 and 4 <= struct_bytes <= LDS_MAX_STRUCT_BYTES


OPS = {"copy": 0, "gather": 1, "scatter": 2}


def auto_variant(op: str, struct_bytes: int) -> str:
    """Name of the variant AUTO resolves to for copy / gather / scatter at this size."""
    v = lib().trove_auto_variant(OPS[op], struct_bytes)
    _check(v if v < 0 else OK, "trove_auto_variant")
    return {c: name for name, c in VARIANTS.items()}[v]


def word_bytes(struct_bytes: int) -> int:
    """Bytes of the word a struct moves in: the widest of 16, 8, 4 dividing it, else 1."""
    for w in (16, 8, 4):
        if struct_bytes % w == 0:
            return w
    return 1

# Struct sizes compiled into the library (bytes).
STRUCT_SIZES = tuple(range(4, 65, 4)) + (96, 128, 160, 192, 224, 256, 384, 512, 768, 1024, 2, 6)


class TroveUnavailable(RuntimeError):
    """libtrove_amd.so is missing or cannot be loaded."""


class TroveError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        msg = strerror(code) if _lib is not None else str(code)
        if code == ERR_HIP and _lib is not None:
            msg += f" (hipError_t {_lib.trove_last_hip_error()})"
        super().__init__(f"{where}: {msg} [{code}]")


_lib = None

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_u64 = ctypes.c_uint64

_SIGNATURES = {
    "trove_strerror": ([_int], ctypes.c_char_p),
    "trove_last_hip_error": ([], _int),
    "trove_version": ([], _int),
    "trove_struct_size_supported": ([_int], _int),
    "trove_struct_word_bytes": ([_int], _int),
    "trove_auto_variant": ([_int, _int], _int),
    "trove_copy": ([_vp, _vp, _i64, _int, _int, _int, _vp], _int),
    "trove_gather": ([_vp, _vp, _vp, _i64, _int, _int, _int, _vp], _int),
    "trove_scatter": ([_vp, _vp, _vp, _i64, _int, _int, _int, _vp], _int),
    "trove_invert_permutation": ([_vp, _vp, _i64, _vp], _int),
    "trove_scatter_permutation": ([_vp, _vp, _vp, _i64, _int, _vp, _vp], _int),
    "trove_scatter_permutation_inverts": ([_int], _int),
    "trove_store_counting": ([_vp, _i64, _int, _int, _int, _vp], _int),
    "trove_load_sum": ([_vp, _vp, _i64, _int, _int, _int, _vp], _int),
    "trove_block_copy": ([_vp, _vp, _i64, _int, _int, _u64, _vp], _int),
    "trove_block_write": ([_vp, _i64, _int, _vp], _int),
    "trove_memcpy_ceiling": ([_vp, _vp, _i64, _vp], _int),
    "trove_random_lines_ceiling": ([_vp, _i64, _vp, _vp, _i64, _vp], _int),
    "trove_random_rows_write_ceiling": ([_vp, _i64, _int, _vp, _i64, _int, _vp], _int),
    "trove_gen_words": ([_vp, _i64, _u64], _int),
    "trove_gen_permutation": ([_vp, _i64, _u64], _int),
    "trove_fill_words_device": ([_vp, _i64, _u64, _vp], _int),
    "trove_gen_indices": ([_vp, _i64, _i64, _int, _u64], _int),
    "trove_gen_zipf": ([_vp, _i64, _i64, ctypes.c_double, _u64], _int),
    "trove_gen_loguniform": ([_vp, _i64, _i64, _i64, _u64], _int),
    "trove_batch_create": ([_vp, ctypes.c_int32, _vp], _int),
    "trove_batch_run": ([_vp, _int, _vp], _int),
    "trove_batch_stats": ([_vp, _vp, _vp], _int),
    "trove_batch_destroy": ([_vp], _int),
    "trove_batch_run_static": ([_vp, ctypes.c_int32, _vp], _int),
    "trove_serve_create": ([_vp, _i64, ctypes.c_int32, _i64, ctypes.c_int32, _i64, ctypes.c_int32, _vp, _vp],
                           _int),
    "trove_serve_submit": ([_vp, _vp, _i64, _int, _vp], _int),
    "trove_serve_step": ([_vp, _vp], _int),
    "trove_serve_poll": ([_vp, _u64, _vp, _vp], _int),
    "trove_serve_output": ([_vp, _u64, _vp], _int),
    "trove_serve_release": ([_vp, _u64], _int),
    "trove_serve_rows": ([_vp, _u64, _vp], _int),
    "trove_serve_fetch": ([_vp, _u64, _vp], _int),
    "trove_serve_drain": ([_vp], _int),
    "trove_serve_stats": ([_vp, _vp, _vp, _vp, _vp, _vp], _int),
    "trove_serve_replay": ([_vp, ctypes.c_int32, _vp, _vp, _vp, _int, _int, _vp, _vp], _int),
    "trove_serve_destroy": ([_vp], _int),
}


def lib():
    """The loaded library (ctypes.CDLL); raises TroveUnavailable if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TroveUnavailable(
                f"{LIB_PATH} is not built; run `python -m trove_amd.build` (or __graft_entry__.build())")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise TroveUnavailable(f"cannot load {LIB_PATH}: {e}") from e
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def declared_functions(header: str = HEADER_PATH) -> list[str]:
    """Names of the functions include/trove_amd.h declares."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(trove_[a-z0-9_]+)\s*\(", text)))


def strerror(code: int) -> str:
    return lib().trove_strerror(code).decode()


def _check(rc: int, where: str) -> None:
    if rc != OK:
        raise TroveError(rc, where)


def _addr(x) -> int:
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError(f"expected a tensor or a device address, got {type(x).__name__}")


def _stream(stream) -> int:
    if stream is None:
        import torch

        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _variant(v) -> int:
    return VARIANTS[v] if isinstance(v, str) else int(v)


def copy(src, dst, n: int, struct_bytes: int, variant="shfl", tile: int = 64, stream=None) -> None:
    """dst[i] = src[i] for n structs (load/store_warp_contiguous<tile> or naive)."""
    _check(lib().trove_copy(_addr(src), _addr(dst), n, struct_bytes, _variant(variant), tile, _stream(stream)),
           "trove_copy")


def gather(src, idx, dst, n: int, struct_bytes: int, variant="tile", tile: int = 64, stream=None) -> None:
    """dst[i] = src[idx[i]] for i < n (tile < 64: coalesced_ptr on tiles of that many lanes)."""
    _check(lib().trove_gather(_addr(src), _addr(idx), _addr(dst), n, struct_bytes, _variant(variant), tile,
                              _stream(stream)), "trove_gather")


def scatter(src, idx, dst, n: int, struct_bytes: int, variant="tile", tile: int = 64, stream=None) -> None:
    """dst[idx[i]] = src[i] for i < n (idx distinct; tile as for gather)."""
    _check(lib().trove_scatter(_addr(src), _addr(idx), _addr(dst), n, struct_bytes, _variant(variant), tile,
                               _stream(stream)), "trove_scatter")


def invert_permutation(perm, inv, n: int, stream=None) -> None:
    """inv[perm[i]] = i for a permutation perm of [0, n)."""
    _check(lib().trove_invert_permutation(_addr(perm), _addr(inv), n, _stream(stream)), "trove_invert_permutation")


def scatter_permutation(src, perm, dst, n: int, struct_bytes: int, workspace=None, stream=None) -> None:
    """dst[perm[i]] = src[i] for a permutation perm of [0, n); with a
    workspace of n int32 it runs as an inversion and a gather where that is
    faster."""
    _check(lib().trove_scatter_permutation(_addr(src), _addr(perm), _addr(dst), n, struct_bytes, _addr(workspace),
                                           _stream(stream)), "trove_scatter_permutation")


def scatter_permutation_inverts(struct_bytes: int) -> bool:
    """Whether scatter_permutation (with a workspace) inverts and gathers at this size."""
    r = lib().trove_scatter_permutation_inverts(struct_bytes)
    _check(r if r < 0 else OK, "trove_scatter_permutation_inverts")
    return bool(r)


def store_counting(dst, n: int, struct_bytes: int, variant="shfl", tile: int = 64, stream=None) -> None:
    """dst[i] = counting_array(i * k) through store_warp_contiguous<tile> (or naive stores)."""
    _check(lib().trove_store_counting(_addr(dst), n, struct_bytes, _variant(variant), tile, _stream(stream)),
           "trove_store_counting")


def load_sum(src, out, n: int, struct_bytes: int, variant="shfl", tile: int = 64, stream=None) -> None:
    """out[i] = wrapping sum of the words of src[i], loaded through load_warp_contiguous<tile>."""
    _check(lib().trove_load_sum(_addr(src), _addr(out), n, struct_bytes, _variant(variant), tile,
                                _stream(stream)), "trove_load_sum")


def block_copy(src, dst, n_elems: int, s: int, divergent: bool = False, seed: int = 0, stream=None) -> None:
    _check(lib().trove_block_copy(_addr(src), _addr(dst), n_elems, s, int(bool(divergent)), seed,
                                  _stream(stream)), "trove_block_copy")


def block_write(dst, n_arrays: int, s: int, stream=None) -> None:
    _check(lib().trove_block_write(_addr(dst), n_arrays, s, _stream(stream)), "trove_block_write")


def memcpy_ceiling(src, dst, nbytes: int, stream=None) -> None:
    _check(lib().trove_memcpy_ceiling(_addr(src), _addr(dst), nbytes, _stream(stream)), "trove_memcpy_ceiling")


def random_lines_ceiling(buf, n_lines: int, line_idx, out, n_reads: int, stream=None) -> None:
    """out[r] = XOR of the 32 words of 128-byte line min(line_idx[r], n_lines - 1) of buf."""
    _check(lib().trove_random_lines_ceiling(_addr(buf), n_lines, _addr(line_idx), _addr(out), n_reads,
                                            _stream(stream)), "trove_random_lines_ceiling")


def random_rows_write_ceiling(buf, n_rows: int, row_bytes: int, row_idx, n_writes: int, nontemporal=False,
                              stream=None) -> None:
    """Row min(row_idx[r], n_rows - 1) of buf (row_bytes in 32, 60, 64, 128)
    receives the words r * (row_bytes / 4) + j, for r < n_writes."""
    _check(lib().trove_random_rows_write_ceiling(_addr(buf), n_rows, row_bytes, _addr(row_idx), n_writes,
                                                 int(bool(nontemporal)), _stream(stream)),
           "trove_random_rows_write_ceiling")


def fill_words_device(out, n: int, seed: int, stream=None) -> None:
    _check(lib().trove_fill_words_device(_addr(out), n, seed, _stream(stream)), "trove_fill_words_device")


def gen_words(n: int, seed: int):
    """Host array of n uint32 equal to what fill_words_device writes."""
    import numpy as np

    out = np.empty(n, dtype=np.uint32)
    _check(lib().trove_gen_words(out.ctypes.data, n, seed), "trove_gen_words")
    return out


def gen_permutation(n: int, seed: int):
    """Host int32 array: a uniform random permutation of [0, n) (Fisher-Yates)."""
    import numpy as np

    out = np.empty(n, dtype=np.int32)
    _check(lib().trove_gen_permutation(out.ctypes.data, n, seed), "trove_gen_permutation")
    return out


PATTERN_IDENTITY = 1
PATTERN_RANDOM = 2
PATTERN_RUNS = 3
PATTERNS = {"identity": PATTERN_IDENTITY, "random": PATTERN_RANDOM, "runs": PATTERN_RUNS}


def gen_indices(n: int, n_src: int, pattern, seed: int):
    """Host int32 indices into [0, n_src): identity, random (a permutation when
    n == n_src) or ascending runs of geometric length (mean 32)."""
    import numpy as np

    pat = PATTERNS[pattern] if isinstance(pattern, str) else int(pattern)
    out = np.empty(n, dtype=np.int32)
    _check(lib().trove_gen_indices(out.ctypes.data, n, n_src, pat, seed), "trove_gen_indices")
    return out


def gen_zipf(n: int, n_items: int, s: float, seed: int):
    """Host int32 rows drawn Zipf(s) over ranks 1..n_items (row = rank - 1)."""
    import numpy as np

    out = np.empty(n, dtype=np.int32)
    _check(lib().trove_gen_zipf(out.ctypes.data, n, n_items, float(s), seed), "trove_gen_zipf")
    return out


def gen_loguniform(n: int, lo: int, hi: int, seed: int):
    import numpy as np

    out = np.empty(n, dtype=np.int64)
    _check(lib().trove_gen_loguniform(out.ctypes.data, n, lo, hi, seed), "trove_gen_loguniform")
    return out


class Job(ctypes.Structure):
    """Mirror of struct trove_job."""
    _fields_ = [("src", ctypes.c_void_p), ("idx", ctypes.c_void_p), ("dst", ctypes.c_void_p),
                ("n", ctypes.c_int64), ("struct_bytes", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("chunk_groups", ctypes.c_int32), ("reserved", ctypes.c_int32)]


BATCH_LAUNCH = 0
BATCH_PERSISTENT = BATCH_LAUNCH  # the same mode's earlier name
BATCH_GRAPH = 1
BATCH_LOOPED = 2
BATCH_RESIDENT = 3
BATCH_MODES = {"launch": BATCH_LAUNCH, "persistent": BATCH_LAUNCH, "graph": BATCH_GRAPH, "looped": BATCH_LOOPED,
               "resident": BATCH_RESIDENT}


class Batch:
    """A validated set of gather and copy jobs runnable as one launch (one
    wavefront per work item, or resident wavefronts striding over them)."""

    def __init__(self, jobs):
        """jobs: iterable of (src, idx, dst, n, struct_bytes, variant, chunk_groups);
        idx None makes the job a contiguous copy."""
        arr = _job_array(jobs)
        self._jobs = arr
        h = ctypes.c_void_p()
        _check(lib().trove_batch_create(ctypes.addressof(arr), len(jobs), ctypes.byref(h)), "trove_batch_create")
        self._h = h

    def run(self, mode="launch", stream=None) -> None:
        m = BATCH_MODES[mode] if isinstance(mode, str) else int(mode)
        _check(lib().trove_batch_run(self._h, m, _stream(stream)), "trove_batch_run")

    def stats(self):
        items, nbytes = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().trove_batch_stats(self._h, ctypes.byref(items), ctypes.byref(nbytes)), "trove_batch_stats")
        return items.value, nbytes.value

    def close(self) -> None:
        if self._h:
            lib().trove_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_STATIC_MAX_JOBS = 4096


def _job_array(jobs):
    arr = (Job * len(jobs))()
    for j, (src, idx, dst, n, b, variant, chunk) in enumerate(jobs):
        arr[j] = Job(_addr(src), _addr(idx), _addr(dst), n, b, _variant(variant), chunk, 0)
    return arr


def batch_run_static(jobs, stream=None) -> None:
    """Runs jobs (tuples as for Batch) in one launch with the job table in the
    library's static device array: no device allocation (trove_batch_run_static)."""
    arr = _job_array(jobs)
    _check(lib().trove_batch_run_static(ctypes.addressof(arr), len(jobs), _stream(stream)), "trove_batch_run_static")


class Server:
    """Continuously batched gather server (trove_serve_*)."""

    def __init__(self, table, n_rows: int, row_bytes: int, budget_bytes: int, max_slots: int = 1024,
                 slice_indices: int = 1 << 16, variant="tile", stream=None):
        h = ctypes.c_void_p()
        self.row_bytes = row_bytes
        self._held = {}  # request id -> device index tensor, until released
        _check(lib().trove_serve_create(_addr(table), n_rows, row_bytes, budget_bytes, max_slots, slice_indices,
                                        _variant(variant), _stream(stream), ctypes.byref(h)), "trove_serve_create")
        self._h = h

    def submit(self, idx, n: int, on_device: bool) -> int:
        """Queues a gather of rows idx[0..n). Host indices: a C-contiguous
        int32 numpy array (copied by the library before this returns).
        Device indices: a contiguous int32 torch tensor on the GPU, kept
        referenced here until the request is released."""
        if n < 0:
            raise ValueError("n must be >= 0")
        if on_device:
            import torch

            if not isinstance(idx, torch.Tensor) or not idx.is_cuda:
                raise ValueError("device indices must be a torch tensor on the GPU")
            if idx.dtype != torch.int32 or not idx.is_contiguous() or idx.numel() < n:
                raise ValueError(f"device indices must be contiguous int32 with at least {n} elements")
            ptr = idx.data_ptr()
        else:
            import numpy as np

            if not isinstance(idx, np.ndarray) or idx.dtype != np.int32 or not idx.flags.c_contiguous:
                raise ValueError("host indices must be a C-contiguous int32 numpy array")
            if idx.size < n:
                raise ValueError(f"host indices hold {idx.size} elements, fewer than n = {n}")
            ptr = idx.ctypes.data
        rid = ctypes.c_uint64()
        _check(lib().trove_serve_submit(self._h, ptr, n, int(on_device), ctypes.byref(rid)), "trove_serve_submit")
        if on_device:
            self._held[rid.value] = idx
        return rid.value

    def step(self) -> int:
        inflight = ctypes.c_int32()
        _check(lib().trove_serve_step(self._h, ctypes.byref(inflight)), "trove_serve_step")
        return inflight.value

    def poll(self, rid: int):
        """(done, latency_s); raises TroveError(ERR_RANGE) for a request
        rejected for a device index outside the table."""
        done, lat = ctypes.c_int32(), ctypes.c_double()
        _check(lib().trove_serve_poll(self._h, rid, ctypes.byref(done), ctypes.byref(lat)), "trove_serve_poll")
        return bool(done.value), lat.value

    def output(self, rid: int) -> int:
        p = ctypes.c_void_p()
        _check(lib().trove_serve_output(self._h, rid, ctypes.byref(p)), "trove_serve_output")
        return p.value or 0

    def rows(self, rid: int) -> int:
        """The request's row count."""
        n = ctypes.c_int64()
        _check(lib().trove_serve_rows(self._h, rid, ctypes.byref(n)), "trove_serve_rows")
        return n.value

    def fetch(self, rid: int, n=None):
        """Host copy (uint8, rows * row_bytes) of a finished request's rows;
        n, when given, must equal the request's row count."""
        import numpy as np

        rows = self.rows(rid)
        if n is not None and n != rows:
            raise ValueError(f"request {rid} has {rows} rows, not {n}")
        out = np.empty(rows * self.row_bytes, dtype=np.uint8)
        _check(lib().trove_serve_fetch(self._h, rid, out.ctypes.data), "trove_serve_fetch")
        return out

    def release(self, rid: int) -> None:
        _check(lib().trove_serve_release(self._h, rid), "trove_serve_release")
        self._held.pop(rid, None)

    def drain(self) -> None:
        _check(lib().trove_serve_drain(self._h), "trove_serve_drain")

    def stats(self):
        c, i, st = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        q, a = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().trove_serve_stats(self._h, *(ctypes.byref(x) for x in (c, i, st, q, a))), "trove_serve_stats")
        return {"completed": c.value, "indices_done": i.value, "steps": st.value, "queued": q.value,
                "active": a.value}

    def replay(self, arrival_s, idx_ptrs, counts, on_device: bool, release_on_done: bool = True,
               return_ids: bool = False):
        """Real-time replay of an arrival trace; returns per-request latency
        (s), or (latency, request ids) with return_ids."""
        import numpy as np

        n = len(counts)
        if len(arrival_s) != n or len(idx_ptrs) != n:
            raise ValueError("arrival_s, idx_ptrs and counts must have the same length")
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        ptrs = np.ascontiguousarray(idx_ptrs, dtype=np.uint64)
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        lat = np.zeros(n, dtype=np.float64)
        ids = np.zeros(n, dtype=np.uint64)
        _check(lib().trove_serve_replay(self._h, n, arr.ctypes.data, ptrs.ctypes.data, cnt.ctypes.data,
                                        int(on_device), int(release_on_done), lat.ctypes.data, ids.ctypes.data),
               "trove_serve_replay")
        return (lat, [int(i) for i in ids]) if return_ids else lat

    def close(self) -> None:
        if self._h:
            lib().trove_serve_destroy(self._h)
            self._h = None
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
