"""ctypes access to the CPU oracle (oracle/oracle.c) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py use
this module, as the checker / CPU reference; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64


REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_parts.so")
REFERENCE_DIR = "/root/reference"


def build() -> str:
    """Compiles oracle.c with the committed Makefile and, when the reference
    sources are present, the reference's pure headers into oracle/_ref/;
    returns the oracle library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir(os.path.join(REFERENCE_DIR, "trove")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REFERENCE={REFERENCE_DIR}"], check=True)
    return LIB_PATH


def ref_lib():
    """The reference-built parts (oracle/_ref), or None when not built."""
    if not os.path.exists(REF_LIB_PATH):
        return None
    L = ctypes.CDLL(REF_LIB_PATH)
    for name in ("ref_gcd_table", "ref_mod_inverse_table", "ref_log_table"):
        getattr(L, name).argtypes = [_i32p]
        getattr(L, name).restype = None
    L.ref_rotate.argtypes = [ctypes.c_int, _i32p, ctypes.c_int, _i32p]
    L.ref_rotate.restype = ctypes.c_int
    L.ref_counting_sum.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p]
    L.ref_counting_sum.restype = ctypes.c_int
    for name in ("ref_mod_inverse_r_table", "ref_bits_table", "ref_inverse_table"):
        getattr(L, name).argtypes = [_i32p]
        getattr(L, name).restype = None
    L.ref_warp_memory.argtypes = [ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p]
    L.ref_warp_memory.restype = ctypes.c_int
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, args in {
            "oracle_copy": [_u8p, _u8p, _i64, ctypes.c_int],
            "oracle_copy_omp": [_u8p, _u8p, _i64, ctypes.c_int],
            "oracle_gather": [_u8p, _i32p, _u8p, _i64, ctypes.c_int],
            "oracle_gather_omp": [_u8p, _i32p, _u8p, _i64, ctypes.c_int],
            "oracle_scatter": [_u8p, _i32p, _u8p, _i64, ctypes.c_int],
            "oracle_counting_store": [_u32p, _i64, ctypes.c_int],
            "oracle_load_sum": [_u32p, _u32p, _i64, ctypes.c_int],
            "oracle_block_copy": [_u32p, _u32p, _i64, ctypes.c_int],
            "oracle_block_write": [_u32p, _i64, ctypes.c_int],
            "oracle_rotate": [_i32p, _i32p, ctypes.c_int, ctypes.c_int],
            "oracle_r2c_tile": [_i32p, ctypes.c_int, ctypes.c_int],
            "oracle_c2r_tile": [_i32p, ctypes.c_int, ctypes.c_int],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int if name.endswith("_tile") else None
        for name in ("oracle_gcd", "oracle_mod_inverse"):
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int]
            getattr(L, name).restype = ctypes.c_int
        L.oracle_mod_inverse_from.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_mod_inverse_from.restype = ctypes.c_int
        L.oracle_pow2_odd_bits.argtypes = [ctypes.c_int]
        L.oracle_pow2_odd_bits.restype = ctypes.c_int
        L.oracle_inverse.argtypes = [_i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_inverse.restype = ctypes.c_int
        L.oracle_warp_load.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u8p]
        L.oracle_warp_store.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, _u8p, ctypes.c_int, ctypes.c_int]
        L.oracle_uncoalesced_store.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, _u8p, ctypes.c_int]
        for name in ("oracle_warp_load", "oracle_warp_store", "oracle_uncoalesced_store"):
            getattr(L, name).restype = None
        L.oracle_floor_log2.argtypes = [ctypes.c_int]
        L.oracle_floor_log2.restype = ctypes.c_int
        L.oracle_schema.argtypes = [ctypes.c_int]
        L.oracle_schema.restype = ctypes.c_int
        L.oracle_max_threads.argtypes = []
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(t)


def copy(src: np.ndarray, n: int, s: int, omp: bool = False) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint8)
    dst = np.zeros(n * s, dtype=np.uint8)
    (lib().oracle_copy_omp if omp else lib().oracle_copy)(_p(src, _u8p), _p(dst, _u8p), n, s)
    return dst


def gather(src: np.ndarray, idx: np.ndarray, s: int, omp: bool = False) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint8)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    n = idx.size
    dst = np.zeros(n * s, dtype=np.uint8)
    fn = lib().oracle_gather_omp if omp else lib().oracle_gather
    fn(_p(src, _u8p), _p(idx, _i32p), _p(dst, _u8p), n, s)
    return dst


def scatter(src: np.ndarray, idx: np.ndarray, dst_init: np.ndarray, s: int) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint8)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    dst = np.array(dst_init, dtype=np.uint8, copy=True)
    lib().oracle_scatter(_p(src, _u8p), _p(idx, _i32p), _p(dst, _u8p), idx.size, s)
    return dst


def counting_store(n: int, k: int) -> np.ndarray:
    dst = np.zeros(n * k, dtype=np.uint32)
    lib().oracle_counting_store(_p(dst, _u32p), n, k)
    return dst


def load_sum(src: np.ndarray, n: int, k: int) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint32)
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_load_sum(_p(src, _u32p), _p(out, _u32p), n, k)
    return out


def block_copy(x: np.ndarray, s: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint32)
    r = np.zeros_like(x)
    lib().oracle_block_copy(_p(x, _u32p), _p(r, _u32p), x.size, s)
    return r


def block_write(l: int, s: int) -> np.ndarray:
    r = np.zeros(l * s, dtype=np.uint32)
    lib().oracle_block_write(_p(r, _u32p), l, s)
    return r


def rotate(t, a: int) -> np.ndarray:
    t = np.ascontiguousarray(t, dtype=np.int32)
    r = np.zeros_like(t)
    lib().oracle_rotate(_p(t, _i32p), _p(r, _i32p), t.size, a)
    return r


def r2c_tile(d: np.ndarray) -> np.ndarray:
    """Reference R2C on a tile given as an (n, m) int32 array of lane slots."""
    d = np.array(d, dtype=np.int32, copy=True)
    n, m = d.shape
    if lib().oracle_r2c_tile(_p(d, _i32p), n, m) != 0:
        raise ValueError("unsupported tile shape")
    return d


def c2r_tile(d: np.ndarray) -> np.ndarray:
    d = np.array(d, dtype=np.int32, copy=True)
    n, m = d.shape
    if lib().oracle_c2r_tile(_p(d, _i32p), n, m) != 0:
        raise ValueError("unsupported tile shape")
    return d


def inverse(fn, x: int, p: int = 0) -> int:
    """First p' >= p with fn[p'] == x (-1 if none), fn an int sequence."""
    fn = np.ascontiguousarray(fn, dtype=np.int32)
    return lib().oracle_inverse(_p(fn, _i32p), fn.size, x, p)


def warp_memory(src: np.ndarray, eb: int, s: int, offset, stride, stored_init: np.ndarray):
    """The reference's strided per-lane access for a tile of lanes: lane l's
    array is warp_load(src, offset[l], stride[l]); returns (loaded, stored,
    uncoalesced): loaded[l] = that array, stored = stored_init after every
    lane's warp_store(array, stored, offset[l], stride[l]) in lane order, and
    uncoalesced[l] = uncoalesced_store(array, row l) (uint8 arrays)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    lanes = len(offset)
    loaded = np.zeros((lanes, s * eb), dtype=np.uint8)
    stored = np.array(stored_init, dtype=np.uint8, copy=True)
    unc = np.zeros((lanes, s * eb), dtype=np.uint8)
    L = lib()
    for l in range(lanes):
        row = loaded[l]
        L.oracle_warp_load(_p(src, _u8p), eb, s, int(offset[l]), int(stride[l]), row.ctypes.data_as(_u8p))
        L.oracle_warp_store(row.ctypes.data_as(_u8p), eb, s, _p(stored, _u8p), int(offset[l]), int(stride[l]))
        urow = unc[l]
        L.oracle_uncoalesced_store(row.ctypes.data_as(_u8p), eb, s, urow.ctypes.data_as(_u8p), 3)
    return loaded, stored, unc
