"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something it does not share code
with: the reference's known answers (tests/golden fixtures), numpy / Python
library routines (memcpy-equivalent copies, fancy indexing, math.gcd,
pow(a, -1, m), np.roll) and closed-form post-conditions of the transposition
(R2C maps lane l slot j = l + j*n to l*m + j; C2R is its inverse), so that a
dropped term, a wrong index or a transposed operand fails here.
"""
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("s", [1, 2, 3, 4, 5, 6, 8, 12, 20, 60, 64, 96, 256])
@pytest.mark.parametrize("n", [0, 1, 63, 64, 1000])
def test_copy_matches_numpy(n, s):
    rng = np.random.default_rng(n * 1000 + s)
    src = rng.integers(0, 256, size=n * s, dtype=np.uint8)
    assert np.array_equal(oracle.copy(src, n, s), src)
    assert np.array_equal(oracle.copy(src, n, s, omp=True), src)


@pytest.mark.parametrize("s", [4, 20, 32, 60, 64])
def test_gather_scatter_match_numpy_indexing(s):
    rng = np.random.default_rng(s)
    n_src, n = 777, 1000
    src = rng.integers(0, 256, size=n_src * s, dtype=np.uint8)
    idx = rng.integers(0, n_src, size=n, dtype=np.int32)  # repeats allowed for gather
    want = src.reshape(n_src, s)[idx].reshape(-1)
    assert np.array_equal(oracle.gather(src, idx, s), want)
    assert np.array_equal(oracle.gather(src, idx, s, omp=True), want)
    perm = rng.permutation(n_src).astype(np.int32)
    init = rng.integers(0, 256, size=n_src * s, dtype=np.uint8)
    want = init.reshape(n_src, s).copy()
    want[perm] = src.reshape(n_src, s)
    assert np.array_equal(oracle.scatter(src, perm, init, s), want.reshape(-1))


def test_gather_scatter_golden():
    g = golden("index_kat.npz")
    n, perm = int(g["n"]), g["perm"]
    for k in range(1, 17):
        src = np.arange(n * k, dtype=np.uint32).view(np.uint8)
        assert np.array_equal(oracle.gather(src, perm, 4 * k).view(np.uint32), g[f"gather_k{k}"])
        out = oracle.scatter(src, perm, np.zeros(n * k * 4, np.uint8), 4 * k)
        assert np.array_equal(out.view(np.uint32), g[f"scatter_k{k}"])


def test_counting_and_sum_golden():
    g = golden("counting_kat.npz")
    n = int(g["n"])
    for k in range(1, 17):
        words = oracle.counting_store(n, k)
        assert np.array_equal(words, g[f"words_k{k}"])
        assert np.array_equal(oracle.load_sum(words, n, k), g[f"sums_k{k}"])


def test_load_sum_wraps_like_uint32():
    rng = np.random.default_rng(3)
    n, k = 500, 13
    src = rng.integers(0, 2**32, size=n * k, dtype=np.uint64).astype(np.uint32)
    want = (src.reshape(n, k).astype(np.uint64).sum(axis=1) % (1 << 32)).astype(np.uint32)
    assert np.array_equal(oracle.load_sum(src, n, k), want)


def test_block_kat_golden():
    g = golden("block_kat.npz")
    l, s = int(g["l"]), int(g["s"])
    written = oracle.block_write(l, s)
    assert np.array_equal(written, g["expected"])
    assert np.array_equal(oracle.block_copy(written, s), g["expected"])


def test_block_kat_reference_size():
    # tests/block.cu:57: l = 100000*256 + 17 arrays of 5 ints.
    l, s = 100000 * 256 + 17, 5
    assert np.array_equal(oracle.block_write(l, s), np.arange(l * s, dtype=np.uint32))


@pytest.mark.parametrize("s", range(1, 17))
def test_block_copy_tail(s):
    rng = np.random.default_rng(s)
    x = rng.integers(0, 2**32, size=64 * s * 3 + 2 * s + (s - 1), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(oracle.block_copy(x, s), x)


def test_number_theory_against_library_routines():
    L = oracle.lib()
    for u in range(1, 130):
        for v in range(1, 130):
            assert L.oracle_gcd(u, v) == math.gcd(u, v), (u, v)
    for m in range(1, 130):
        for a in range(0, 200):
            try:
                want = pow(a, -1, m) if m > 1 else -1
                if want == 0:
                    want = -1
            except ValueError:
                want = -1
            assert L.oracle_mod_inverse(a, m) == want, (a, m)
    for m in range(1, 5000):
        assert L.oracle_floor_log2(m) == m.bit_length() - 1


@pytest.mark.parametrize("m", range(1, 40))
def test_rotate_is_left_roll(m):
    t = np.arange(100, 100 + m, dtype=np.int32)
    limit = 1 << max(1, (m - 1).bit_length())
    for a in range(limit):
        assert np.array_equal(oracle.rotate(t, a), np.roll(t, -(a % m))), (m, a)


def test_schema_classification():
    L = oracle.lib()
    # transpose.h:38-55: power of two -> po2 (1), odd -> odd (0), else composite (2)
    assert [L.oracle_schema(m) for m in (2, 3, 4, 5, 6, 8, 10, 12, 15, 16)] == [1, 0, 1, 0, 2, 1, 2, 2, 0, 1]


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64])
def test_reference_schedule_post_conditions(n):
    for m in range(2, n + 1):
        col = np.array([[l + j * n for j in range(m)] for l in range(n)], dtype=np.int32)
        row = np.array([[l * m + j for j in range(m)] for l in range(n)], dtype=np.int32)
        assert np.array_equal(oracle.r2c_tile(col), row), ("r2c", n, m)
        assert np.array_equal(oracle.c2r_tile(row), col), ("c2r", n, m)


def test_schedule_check_detects_a_planted_mistake():
    """A tile whose lanes are swapped must fail the post-condition check."""
    n, m = 32, 5
    col = np.array([[l + j * n for j in range(m)] for l in range(n)], dtype=np.int32)
    wrong = col.copy()
    wrong[[3, 4]] = wrong[[4, 3]]
    row = np.array([[l * m + j for j in range(m)] for l in range(n)], dtype=np.int32)
    assert not np.array_equal(oracle.r2c_tile(wrong), row)


# ---- against the reference's own code (oracle/_ref, built from the
# reference headers in place where they compile: number theory, rotate,
# counting_array / sum) ----

def _ref():
    L = oracle.ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (the reference sources are not present here)")
    return L


def _tab(fn, shape):
    out = np.zeros(shape, dtype=np.int32)
    fn(out.ctypes.data_as(oracle._i32p))
    return out


def test_gcd_and_inverse_match_reference_build():
    R, L = _ref(), oracle.lib()
    g = _tab(R.ref_gcd_table, 33 * 33).reshape(33, 33)
    inv = _tab(R.ref_mod_inverse_table, 33 * 33).reshape(33, 33)
    lg = _tab(R.ref_log_table, 33)
    for u in range(1, 33):
        for v in range(1, 33):
            assert L.oracle_gcd(u, v) == g[u, v], (u, v)
    for a in range(0, 33):
        for m in range(1, 33):
            assert L.oracle_mod_inverse(a, m) == inv[a, m], (a, m)
    for m in range(1, 33):
        assert L.oracle_floor_log2(m) == lg[m]


@pytest.mark.parametrize("m", range(1, 17))
def test_rotate_matches_reference_build(m):
    R = _ref()
    t = np.arange(7, 7 + m, dtype=np.int32)
    limit = 1 << max(1, (m - 1).bit_length())
    for a in range(limit):
        r = np.zeros(m, dtype=np.int32)
        assert R.ref_rotate(m, t.ctypes.data_as(oracle._i32p), a, r.ctypes.data_as(oracle._i32p)) == 0
        assert np.array_equal(oracle.rotate(t, a), r), (m, a)


def test_counting_and_sum_match_reference_build():
    R = _ref()
    n = 64 * 2 + 17
    for k in range(1, 17):
        words = np.zeros(k, dtype=np.int32)
        mine = oracle.counting_store(n, k)
        sums = oracle.load_sum(mine, n, k)
        for g in (0, 1, 63, 64, n - 1):
            s = R.ref_counting_sum(k, g * k, 1, words.ctypes.data_as(oracle._i32p))
            assert np.array_equal(words.view(np.uint32), mine[g * k:(g + 1) * k]), (k, g)
            assert np.uint32(s & 0xFFFFFFFF) == sums[g], (k, g)
