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


def test_inverse_from_start_and_traits_closed_forms():
    L = oracle.lib()
    for m in range(1, 70):
        for a in range(0, 70):
            for r in range(1, 8):
                want = next((x for x in range(r, m) if (x * a) % m == 1), -1)
                assert L.oracle_mod_inverse_from(a, m, r) == want, (a, m, r)
            assert L.oracle_mod_inverse_from(a, m, 1) == L.oracle_mod_inverse(a, m)
    for m in range(0, 2000):
        pow2 = m == 0 or bin(m).count("1") == 1  # the reference's (m & (m - 1)) == 0 holds for 0 too
        assert L.oracle_pow2_odd_bits(m) == int(pow2) + 2 * (m % 2), m
    fn = [(5 * p + 3) % 64 for p in range(200)]
    for x in range(64):
        for p0 in (0, 10, 63, 64):
            want = next(p for p in range(p0, 200) if fn[p] == x)
            assert oracle.inverse(fn, x, p0) == want, (x, p0)
    assert oracle.inverse(fn, 64, 0) == -1  # outside the map's range


def _memory_case(rng, s, eb, lanes, coalesced):
    """Offsets / strides with distinct positions across lanes (disjoint
    segments per lane, any stride >= 1 inside), or the coalesced pattern
    offset = lane, stride = lanes."""
    if coalesced:
        offset = np.arange(lanes)
        stride = np.full(lanes, lanes)
        span = lanes * s
    else:
        seg = 4 * s + 3
        stride = rng.integers(1, 5, size=lanes)
        lead = np.array([rng.integers(0, seg - (s - 1) * int(st)) for st in stride])
        offset = rng.permutation(lanes) * seg + lead
        span = lanes * seg
    return offset.astype(np.int32), stride.astype(np.int32), span


@pytest.mark.parametrize("s,eb", [(1, 4), (2, 4), (5, 4), (16, 4), (3, 8), (2, 16), (5, 16)])
def test_warp_memory_closed_form(s, eb):
    rng = np.random.default_rng(s * 100 + eb)
    for coalesced in (True, False):
        offset, stride, span = _memory_case(rng, s, eb, 64, coalesced)
        src = rng.integers(0, 256, size=span * eb, dtype=np.uint8)
        init = rng.integers(0, 256, size=span * eb, dtype=np.uint8)
        loaded, stored, unc = oracle.warp_memory(src, eb, s, offset, stride, init)
        el = src.reshape(span, eb)
        want_stored = init.reshape(span, eb).copy()
        for l in range(64):
            pos = offset[l] + stride[l] * np.arange(s)
            assert np.array_equal(loaded[l].reshape(s, eb), el[pos]), (l, "warp_load")
            want_stored[pos] = el[pos]
        assert np.array_equal(stored, want_stored.reshape(-1)), "warp_store"
        assert np.array_equal(unc, loaded), "uncoalesced_store ignores its stride"
    # A planted mistake: a lane whose stride is off by one must load other words.
    # (Buffers padded by s words so the longer stride stays in bounds.)
    offset, stride, span = _memory_case(rng, s, eb, 64, False)
    pad = span + 64 * s
    src = rng.integers(0, 256, size=pad * eb, dtype=np.uint8)
    if s > 1:
        wrong, _, _ = oracle.warp_memory(src, eb, s, offset, stride + 1, np.zeros(pad * eb, np.uint8))
        good, _, _ = oracle.warp_memory(src, eb, s, offset, stride, np.zeros(pad * eb, np.uint8))
        assert not np.array_equal(wrong, good)


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


def test_inverse_from_start_and_traits_match_reference_build():
    R, L = _ref(), oracle.lib()
    inv_r = _tab(R.ref_mod_inverse_r_table, 6 * 33 * 32).reshape(6, 33, 32)
    for r in range(1, 7):
        for a in range(0, 33):
            for m in range(r, 33):
                assert L.oracle_mod_inverse_from(a, m, r) == inv_r[r - 1, a, m - 1], (a, m, r)
    bits = _tab(R.ref_bits_table, 257)
    for m in range(0, 257):
        assert L.oracle_pow2_odd_bits(m) == bits[m], m
    inv = _tab(R.ref_inverse_table, 128)
    fn = [(5 * p + 3) % 64 for p in range(200)]
    for x in range(64):
        assert oracle.inverse(fn, x, 0) == inv[x], x
        assert oracle.inverse(fn, x, 10) == inv[64 + x], x


@pytest.mark.parametrize("s", [1, 2, 5, 16])
def test_warp_memory_matches_reference_build(s):
    R = _ref()
    rng = np.random.default_rng(900 + s)
    for coalesced in (True, False):
        offset, stride, span = _memory_case(rng, s, 4, 64, coalesced)
        src = rng.integers(-2**31, 2**31, size=span, dtype=np.int64).astype(np.int32)
        init = rng.integers(-2**31, 2**31, size=span, dtype=np.int64).astype(np.int32)
        loaded, stored, unc = oracle.warp_memory(src.view(np.uint8), 4, s, offset, stride, init.view(np.uint8))
        ref_stored = init.copy()
        for l in range(64):
            rl = np.zeros(s, dtype=np.int32)
            ru = np.zeros(s, dtype=np.int32)
            assert R.ref_warp_memory(s, src.ctypes.data_as(oracle._i32p), int(offset[l]), int(stride[l]),
                                     rl.ctypes.data_as(oracle._i32p), ref_stored.ctypes.data_as(oracle._i32p),
                                     ru.ctypes.data_as(oracle._i32p)) == 0
            assert np.array_equal(loaded[l].view(np.int32), rl), (s, l)
            assert np.array_equal(unc[l].view(np.int32), ru), (s, l)
        assert np.array_equal(stored.view(np.int32), ref_stored), s
