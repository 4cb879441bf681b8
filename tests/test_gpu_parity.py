"""GPU parity: every entry point of libtrove_amd against the CPU oracle on the
same seeded inputs, bit-exact, through the C ABI.

Sizes keep the oracle fast; the reference's full-size workloads are covered by
size-independent properties (identity copy, gather against torch indexing and
against the oracle on sampled rows). Partial last wavefronts (n % 64 != 0)
exercise the coalesced_ptr tile cascade (32/16/8/4/2 lanes and the per-lane
fallback); the divergent block test exercises warp_converged() == false.
"""
import os

import numpy as np
import pytest
import torch

import oracle
import trove_amd as T

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = T.STRUCT_SIZES
TILES = (64, 32, 16, 8, 4, 2)


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rand_bytes(nbytes: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(0, 256, size=nbytes, dtype=np.uint8)


@pytest.mark.parametrize("s", SIZES)
def test_copy_all_variants(trove, cuda, s):
    n = 64 * 37 + 13
    src = rand_bytes(n * s, s)
    want = oracle.copy(src, n, s)
    d_src = dev(src)
    cases = (("shfl", TILES), ("direct", (64,)), ("direct_launch", (64,))) + \
        ((("lds", (64,)),) if T.lds_supported(s) else ())
    for variant, tiles in cases:
        for tile in tiles:
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.copy(d_src, d_dst, n, s, variant=variant, tile=tile)
            assert np.array_equal(host(d_dst), want), (s, variant, tile)
    if not T.lds_supported(s):
        with pytest.raises(T.TroveError) as e:
            T.copy(d_src, d_src, n, s, variant="lds")
        assert e.value.code == T.ERR_VARIANT


def test_seeded_random_cases(trove, cuda):
    # 120 cases drawn from one seed: struct size, operation, variant (and tile
    # for copies), length log-uniform over 1..8000, gather indices with
    # repeats, scatters of distinct indices into a larger pre-filled
    # destination; each bit-exact against the oracle. The case is in the
    # assertion message.
    rng = np.random.default_rng(20240611)
    for case in range(120):
        s = int(rng.choice(SIZES))
        op = str(rng.choice(["copy", "gather", "scatter"]))
        n = int(np.exp(rng.uniform(0, np.log(8000))))
        lds = ("lds",) if T.lds_supported(s) else ()
        tile = 64
        if op == "copy":
            variant = str(rng.choice(("shfl", "direct", "direct_launch", "auto") + lds))
            if variant == "shfl":
                tile = int(rng.choice(TILES))
            elif variant == "auto":
                tile = int(rng.choice((64, 32)))
        else:
            variant = str(rng.choice(("shfl", "tile", "direct", "direct_launch", "auto") + lds))
        what = (case, op, s, n, variant, tile)
        src = rand_bytes(n * s, 5000 + case)
        if op == "copy":
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.copy(dev(src), d_dst, n, s, variant=variant, tile=tile)
            assert np.array_equal(host(d_dst), oracle.copy(src, n, s)), what
        elif op == "gather":
            n_src = int(rng.integers(1, 3 * n + 2))
            table = rand_bytes(n_src * s, 6000 + case)
            idx = rng.integers(0, n_src, size=n, dtype=np.int32)
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.gather(dev(table), dev(idx), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst), oracle.gather(table, idx, s)), what
        else:
            n_dst = n + int(rng.integers(0, 2 * n + 1))
            idx = rng.permutation(n_dst)[:n].astype(np.int32)
            init = rand_bytes(n_dst * s, 7000 + case)
            d_dst = dev(init.copy())
            T.scatter(dev(src), dev(idx), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst), oracle.scatter(src, idx, init, s)), what


@pytest.mark.parametrize("s", [4, 12, 20, 32, 60, 64, 96, 160])
def test_copy_partial_wavefront_cascade(trove, cuda, s):
    # Every trailing-group length 1..64 runs the tile cascade once.
    for n in list(range(1, 66)) + [127, 129]:
        src = rand_bytes(n * s, 1000 + n)
        for variant in ("shfl", "lds"):
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.copy(dev(src), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst), src), (s, n, variant)


def _patterns(n, n_src, seed):
    rng = np.random.default_rng(seed)
    runs = []
    while sum(len(r) for r in runs) < n:
        length = int(rng.geometric(1 / 32))
        start = int(rng.integers(0, max(1, n_src - length)))
        runs.append(np.arange(start, min(n_src, start + length)))
    return {
        "identity": np.arange(n, dtype=np.int32) % n_src,
        "random": rng.integers(0, n_src, size=n, dtype=np.int32),
        "runs": np.concatenate(runs)[:n].astype(np.int32),
        "reversed": (n_src - 1 - np.arange(n) % n_src).astype(np.int32),
    }


@pytest.mark.parametrize("s", SIZES)
def test_gather_all_variants(trove, cuda, s):
    n, n_src = 64 * 29 + 37, 3001
    src = rand_bytes(n_src * s, s + 7)
    d_src = dev(src)
    for name, idx in _patterns(n, n_src, s).items():
        want = oracle.gather(src, idx, s)
        d_idx = dev(idx)
        for variant in ("shfl", "tile", "direct", "direct_launch") + (("lds",) if T.lds_supported(s) else ()):
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.gather(d_src, d_idx, d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst), want), (s, name, variant)


@pytest.mark.parametrize("s", SIZES)
def test_scatter_all_variants(trove, cuda, s):
    n = 64 * 23 + 5
    src = rand_bytes(n * s, s + 11)
    init = rand_bytes(n * s, s + 12)
    d_src = dev(src)
    for name, idx in (("perm", np.random.default_rng(s).permutation(n).astype(np.int32)),
                      ("identity", np.arange(n, dtype=np.int32))):
        want = oracle.scatter(src, idx, init, s)
        for variant in ("shfl", "tile", "direct", "direct_launch") + (("lds",) if T.lds_supported(s) else ()):
            d_dst = dev(init.copy())
            T.scatter(d_src, dev(idx), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst), want), (s, name, variant)


@pytest.mark.parametrize("s", [8, 12, 20, 32, 60, 64, 96, 256, 1024, 6])
def test_gather_scatter_sub_wavefront_tiles(trove, cuda, s):
    # trove_gather / trove_scatter with tile 32..2: coalesced_ptr on tiles of
    # that many lanes (AUTO resolves to it), a partial last group included;
    # DIRECT and DIRECT_LAUNCH ignore the width.
    n = 64 * 29 + 37
    rng = np.random.default_rng(900 + s)
    src = rand_bytes(n * s, 910 + s)
    perm = rng.permutation(n).astype(np.int32)
    want_g = oracle.gather(src, perm, s)
    want_s = oracle.scatter(src, perm, np.zeros(n * s, dtype=np.uint8), s)
    d_src, d_perm = dev(src), dev(perm)
    for tile in (32, 16, 8, 4, 2):
        for variant in ("shfl", "auto", "direct", "direct_launch"):
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.gather(d_src, d_perm, d_dst, n, s, variant=variant, tile=tile)
            assert np.array_equal(host(d_dst), want_g), ("gather", s, tile, variant)
            d_dst.zero_()
            T.scatter(d_src, d_perm, d_dst, n, s, variant=variant, tile=tile)
            assert np.array_equal(host(d_dst), want_s), ("scatter", s, tile, variant)


def test_invert_permutation(trove, cuda):
    for n in (1, 63, 64, 65, 64 * 1000 + 7):
        perm = np.random.default_rng(n).permutation(n).astype(np.int32)
        inv = torch.full((n,), -1, dtype=torch.int32, device=cuda)
        T.invert_permutation(dev(perm), inv, n)
        assert np.array_equal(host(inv), np.argsort(perm).astype(np.int32)), n


def test_invert_skips_entries_outside_the_range(trove, cuda):
    # Entries outside [0, n) are skipped and unnamed inverse entries keep
    # their contents (the kernel's own bounds check; nothing is written
    # outside inv).
    n = 64 * 50 + 9
    rng = np.random.default_rng(5)
    perm = rng.permutation(n).astype(np.int32)
    bad = rng.choice(n, size=200, replace=False)
    perm[bad[:100]] = n + rng.integers(0, 1 << 20, size=100)
    perm[bad[100:]] = -1 - rng.integers(0, 1 << 20, size=100)
    inv = torch.full((n,), -7, dtype=torch.int32, device=cuda)
    T.invert_permutation(dev(perm), inv, n)
    want = np.full(n, -7, dtype=np.int32)
    ok = (perm >= 0) & (perm < n)
    want[perm[ok]] = np.nonzero(ok)[0]
    assert np.array_equal(host(inv), want)


@pytest.mark.parametrize("s", [12, 60])
def test_scatter_permutation_of_a_non_permutation_stays_in_bounds(trove, cuda, s):
    # Not a permutation (duplicates and entries outside [0, n)): the result is
    # unspecified, but the workspace is zeroed before the inversion, so a row
    # nothing maps to gathers row 0 and a row one in-range entry maps to
    # gathers that entry's source row.
    k = s // 4
    n = 64 * 40 + 13
    rng = np.random.default_rng(s)
    perm = rng.permutation(n).astype(np.int32)
    perm[rng.choice(n, size=150, replace=False)] = perm[rng.choice(n, size=150, replace=False)]
    perm[rng.choice(n, size=50, replace=False)] = n + 3
    words = T.gen_words(n * k, 17 + s)
    src = dev(words.view(np.int32))
    dst = torch.full((n * k,), -1, dtype=torch.int32, device=cuda)
    ws = torch.full((n,), 12345, dtype=torch.int32, device=cuda)
    assert T.scatter_permutation_inverts(s)
    T.scatter_permutation(src, dev(perm), dst, n, s, workspace=ws)
    got = host(dst).reshape(n, k)
    rows = words.view(np.int32).reshape(n, k)
    ok = (perm >= 0) & (perm < n)
    counts = np.bincount(perm[ok], minlength=n)
    single = np.nonzero(counts == 1)[0]
    first = np.full(n, -1, dtype=np.int64)
    first[perm[ok][::-1]] = np.nonzero(ok)[0][::-1]
    assert np.array_equal(got[single], rows[first[single]])
    assert np.array_equal(got[counts == 0], np.broadcast_to(rows[0], (int((counts == 0).sum()), k)))
    for j in np.nonzero(counts > 1)[0]:
        assert any(np.array_equal(got[j], rows[i]) for i in np.nonzero(perm == j)[0]), j


@pytest.mark.parametrize("s", SIZES)
def test_scatter_permutation_all_sizes(trove, cuda, s):
    # dst[perm[i]] = src[i]: the oracle's scatter, through the inversion and
    # a gather (12..60 bytes) or the plain scatter (other sizes, or no
    # workspace); partial last group included.
    n = 64 * 37 + 11
    rng = np.random.default_rng(700 + s)
    perm = rng.permutation(n).astype(np.int32)
    src = rand_bytes(n * s, 800 + s)
    want = oracle.scatter(src, perm, np.zeros(n * s, dtype=np.uint8), s)
    d_src, d_perm = dev(src), dev(perm)
    ws = torch.empty(n, dtype=torch.int32, device=cuda)
    for workspace in (ws, None):
        d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
        T.scatter_permutation(d_src, d_perm, d_dst, n, s, workspace=workspace)
        assert np.array_equal(host(d_dst), want), (s, workspace is None)


@pytest.mark.parametrize("s", [4, 8, 12, 16, 20, 60, 64, 96, 256])
def test_gather_scatter_fewer_than_a_group(trove, cuda, s):
    # n below, at and just past one 64-struct group: every variant's partial
    # group handling, against the oracle.
    n_src = 200
    src = rand_bytes(n_src * s, 60 + s)
    d_src = dev(src)
    variants = ("shfl", "tile", "direct", "direct_launch") + (("lds",) if T.lds_supported(s) else ())
    for n in (1, 5, 63, 64, 65, 129):
        idx = np.random.default_rng(n + s).integers(0, n_src, size=n, dtype=np.int32)
        perm = np.random.default_rng(n + s + 1).permutation(n).astype(np.int32)
        init = rand_bytes(n * s, 61 + s)
        for variant in variants + ("auto",):
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.gather(d_src, dev(idx), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst), oracle.gather(src, idx, s)), (s, n, variant, "gather")
            d_dst = dev(init.copy())
            T.scatter(d_src, dev(perm), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst), oracle.scatter(src[:n * s], perm, init, s)), (s, n, variant, "scatter")


def test_auto_variant_every_size(trove, cuda):
    # AUTO picks a path per struct size (LDS-staged or not); whatever it picks
    # must match the oracle, and a sub-wavefront tile with AUTO must work.
    n, n_src = 64 * 11 + 29, 1500
    for s in SIZES:
        src = rand_bytes(n_src * s, 40 + s)
        d_src = dev(src)
        for tile in (64, 32):
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.copy(d_src, d_dst, n, s, variant="auto", tile=tile)
            assert np.array_equal(host(d_dst), src[:n * s]), (s, "copy", tile)
        idx = np.random.default_rng(s).integers(0, n_src, size=n, dtype=np.int32)
        d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
        T.gather(d_src, dev(idx), d_dst, n, s, variant="auto")
        assert np.array_equal(host(d_dst), oracle.gather(src, idx, s)), (s, "gather")
        perm = np.random.default_rng(s + 1).permutation(n).astype(np.int32)
        init = rand_bytes(n * s, 41 + s)
        d_dst = dev(init.copy())
        T.scatter(d_src, dev(perm), d_dst, n, s, variant="auto")
        assert np.array_equal(host(d_dst), oracle.scatter(src[:n * s], perm, init, s)), (s, "scatter")


@pytest.mark.parametrize("s", [4, 8, 12, 20, 24, 40, 44, 56, 60])
def test_lds_unaligned_tile_bases(trove, cuda, s):
    # Tile bases that are word-aligned but not 16-byte aligned take the LDS
    # variant's dword-chunk path: copy (both sides), gather (destination) and
    # scatter (source) at byte offsets of one and two dwords, against the
    # oracle; bytes outside the views stay untouched.
    w = T.word_bytes(s)
    n, n_src = 64 * 19 + 21, 2500
    src = rand_bytes(n_src * s, s + 3)
    idx = np.random.default_rng(s).integers(0, n_src, size=n, dtype=np.int32)
    perm = np.random.default_rng(s + 1).permutation(n).astype(np.int32)
    for off in (4, 8):
        if off % w:
            continue
        pad = rand_bytes(n_src * s + 16, off)
        d_src = dev(pad.copy())
        d_src[off:off + n_src * s] = dev(src)
        d_out = dev(pad.copy())
        T.copy(d_src[off:], d_out[off:], n, s, variant="lds")
        want = pad.copy()
        want[off:off + n * s] = src[:n * s]
        assert np.array_equal(host(d_out), want), (s, off, "copy")
        d_out = dev(pad.copy())
        T.gather(d_src[off:], dev(idx), d_out[off:], n, s, variant="lds")
        want = pad.copy()
        want[off:off + n * s] = oracle.gather(src, idx, s)
        assert np.array_equal(host(d_out), want), (s, off, "gather")
        d_out = dev(pad.copy())
        T.scatter(d_src[off:], dev(perm), d_out[off:], n, s, variant="lds")
        want = pad.copy()
        want[off:off + n * s] = oracle.scatter(src[:n * s], perm, pad[off:off + n * s], s)
        assert np.array_equal(host(d_out), want), (s, off, "scatter")


def test_index_golden(trove, cuda):
    g = np.load(os.path.join(GOLDEN, "index_kat.npz"), allow_pickle=False)
    n, perm = int(g["n"]), g["perm"]
    for k in range(1, 17):
        s = 4 * k
        d_src = dev(np.arange(n * k, dtype=np.uint32).view(np.uint8))
        for variant in ("shfl", "tile", "direct", "lds"):
            d_dst = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
            T.gather(d_src, dev(perm), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst).view(np.uint32), g[f"gather_k{k}"]), (k, variant)
            d_dst.zero_()
            T.scatter(d_src, dev(perm), d_dst, n, s, variant=variant)
            assert np.array_equal(host(d_dst).view(np.uint32), g[f"scatter_k{k}"]), (k, variant)


def test_counting_store_and_load_sum(trove, cuda):
    g = np.load(os.path.join(GOLDEN, "counting_kat.npz"), allow_pickle=False)
    n_g = int(g["n"])
    for k in range(1, 17):
        s = 4 * k
        for n in (n_g, 64 * 41 + 9):
            want_words = oracle.counting_store(n, k)
            for variant in ("shfl", "direct"):
                d = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
                T.store_counting(d, n, s, variant=variant)
                words = host(d).view(np.uint32)
                assert np.array_equal(words, want_words), (k, n, variant)
                if n == n_g:
                    assert np.array_equal(words, g[f"words_k{k}"])
                src = np.random.default_rng(k).integers(0, 2**32, size=n * k, dtype=np.uint64).astype(np.uint32)
                out = torch.zeros(n, dtype=torch.int32, device=cuda)
                T.load_sum(dev(src.view(np.uint8)), out, n, s, variant=variant)
                assert np.array_equal(host(out).view(np.uint32), oracle.load_sum(src, n, k)), (k, n, variant)


@pytest.mark.parametrize("tile", [32, 16, 8, 4, 2])
def test_counting_store_and_load_sum_tiles(trove, cuda, tile):
    # The reference program's TILE_SIZE (tests/Makefile:4, 32 by default):
    # store_warp_contiguous / load_warp_contiguous on sub-wavefront tiles,
    # including structs of more words than the tile (the direct path there).
    n = 64 * 37 + 29
    for k in (1, 2, 3, 4, 5, 8, 15, 16, 24, 64):
        s = 4 * k
        d = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
        T.store_counting(d, n, s, variant="shfl", tile=tile)
        assert np.array_equal(host(d).view(np.uint32), oracle.counting_store(n, k)), (k, tile)
        src = np.random.default_rng(k * 7 + tile).integers(0, 2**32, size=n * k, dtype=np.uint64).astype(np.uint32)
        out = torch.zeros(n, dtype=torch.int32, device=cuda)
        T.load_sum(dev(src.view(np.uint8)), out, n, s, variant="shfl", tile=tile)
        assert np.array_equal(host(out).view(np.uint32), oracle.load_sum(src, n, k)), (k, tile)


@pytest.mark.parametrize("k", [1, 3, 4, 5, 8, 10, 13, 16, 24, 40, 64])
def test_counting_store_and_load_sum_many_groups(trove, cuda, k):
    # Enough 64-struct groups that the launch spreads work-groups over all
    # XCDs in XCD-contiguous order under a residency cap, with a partial
    # last group; sizes up to 256 bytes.
    s = 4 * k
    n = (1 << 20) + 37 if k <= 16 else (1 << 18) + 37
    want_words = oracle.counting_store(n, k)
    src = np.random.default_rng(100 + k).integers(0, 2**32, size=n * k, dtype=np.uint64).astype(np.uint32)
    want_sum = oracle.load_sum(src, n, k)
    d_src = dev(src.view(np.uint8))
    for variant in ("shfl", "direct", "auto"):
        d = torch.zeros(n * s, dtype=torch.uint8, device=cuda)
        T.store_counting(d, n, s, variant=variant)
        assert np.array_equal(host(d).view(np.uint32), want_words), (k, variant)
        out = torch.zeros(n, dtype=torch.int32, device=cuda)
        T.load_sum(d_src, out, n, s, variant=variant)
        assert np.array_equal(host(out).view(np.uint32), want_sum), (k, variant)


@pytest.mark.parametrize("s", range(1, 17))
def test_block_copy(trove, cuda, s):
    e = 64 * s * 53 + 3 * s + (s - 1)
    x = np.random.default_rng(s).integers(0, 2**32, size=e, dtype=np.uint64).astype(np.uint32)
    # float32 raw bit patterns: NaNs with payloads, infinities, denormals, -0.
    x[:6] = np.array([0x7FC00001, 0xFF800000, 0x7F800000, 0x00000001, 0x80000000, 0x7FBFFFFF], dtype=np.uint32)
    want = oracle.block_copy(x, s)
    d_x = dev(x.view(np.int32))
    for divergent in (False, True):
        d_r = torch.zeros(e, dtype=torch.int32, device=cuda)
        T.block_copy(d_x, d_r, e, s, divergent=divergent, seed=1234 + s)
        assert np.array_equal(host(d_r).view(np.uint32), want), (s, divergent)


def test_block_write_kat(trove, cuda):
    g = np.load(os.path.join(GOLDEN, "block_kat.npz"), allow_pickle=False)
    l, s = int(g["l"]), int(g["s"])
    d = torch.zeros(l * s, dtype=torch.int32, device=cuda)
    T.block_write(d, l, s)
    assert np.array_equal(host(d).view(np.uint32), g["expected"])
    # The reference's own size (tests/block.cu:57): 100000*256 + 17 arrays.
    l = 100000 * 256 + 17
    d = torch.zeros(l * s, dtype=torch.int32, device=cuda)
    T.block_write(d, l, s)
    want = torch.arange(l * s, dtype=torch.int64, device=cuda).to(torch.int32)
    assert torch.equal(d, want)
    r = torch.zeros_like(d)
    T.block_copy(d, r, l * s, s)
    assert torch.equal(r, want)


def test_memcpy_ceiling_and_fill(trove, cuda):
    n = (1 << 20) + 64
    d = torch.zeros(n, dtype=torch.int32, device=cuda)
    T.fill_words_device(d, n, 99)
    assert np.array_equal(host(d).view(np.uint32), T.gen_words(n, 99))
    r = torch.zeros_like(d)
    T.memcpy_ceiling(d, r, n * 4)
    assert torch.equal(r, d)


@pytest.mark.parametrize("n_reads", [1, 255, 256, 257, 10007])
def test_random_lines_ceiling_values(trove, cuda, n_reads):
    # out[r] = XOR of the 32 words of line min(idx[r], n_lines - 1) (indices
    # beyond the buffer read its last line instead of outside it).
    n_lines = 4096
    words = T.gen_words(n_lines * 32, 5)
    rng = np.random.default_rng(n_reads)
    idx = rng.integers(0, 3 * n_lines, size=n_reads, dtype=np.int64).astype(np.uint32)
    d_buf = dev(words.view(np.int32))
    out = torch.zeros(n_reads, dtype=torch.int32, device=cuda)
    T.random_lines_ceiling(d_buf, n_lines, dev(idx.view(np.int32)), out, n_reads)
    want = np.bitwise_xor.reduce(words.reshape(n_lines, 32)[np.minimum(idx, n_lines - 1)], axis=1)
    assert np.array_equal(host(out).view(np.uint32), want)


@pytest.mark.parametrize("row_bytes", [32, 60, 64, 128])
def test_random_rows_write_ceiling_values(trove, cuda, row_bytes):
    # Write r lands at row idx[r] with the words r * k + j; untouched rows
    # keep their contents (both store policies; partial last rounds).
    k = row_bytes // 4
    n_rows = 64 * 97 + 13
    rng = np.random.default_rng(row_bytes)
    for nt in (False, True):
        for n_writes in (1, 100, n_rows):
            idx = rng.permutation(n_rows)[:n_writes].astype(np.int32)
            init = rng.integers(0, 2**32, size=n_rows * k, dtype=np.uint64).astype(np.uint32)
            d_buf = dev(init.view(np.int32))
            T.random_rows_write_ceiling(d_buf, n_rows, row_bytes, dev(idx), n_writes, nontemporal=nt)
            want = init.reshape(n_rows, k).copy()
            want[idx] = (np.arange(n_writes, dtype=np.uint64)[:, None] * k + np.arange(k)).astype(np.uint32)
            assert np.array_equal(host(d_buf).view(np.uint32).reshape(n_rows, k), want), (nt, n_writes)


def test_small_contiguous_config_bit_exact(trove, cuda):
    # Contiguous AoS copy, 20-byte structs, N = 2^20, against the scalar loop.
    n, s = 1 << 20, 20
    words = T.gen_words(n * s // 4, 2024)
    d_src = torch.zeros(n * s // 4, dtype=torch.int32, device=cuda)
    T.fill_words_device(d_src, n * s // 4, 2024)
    d_dst = torch.zeros_like(d_src)
    T.copy(d_src, d_dst, n, s, variant="shfl")
    want = oracle.copy(words.view(np.uint8), n, s)
    assert np.array_equal(host(d_dst).view(np.uint8), want)


def _equal_in_chunks(a: torch.Tensor, b: torch.Tensor, chunk: int = 1 << 26) -> bool:
    # torch's own comparison kernels reject launches over ~2^29 elements.
    a, b = a.reshape(-1), b.reshape(-1)
    if a.numel() != b.numel():
        return False
    return all(torch.equal(a[i:i + chunk], b[i:i + chunk]) for i in range(0, a.numel(), chunk))


@pytest.fixture(scope="module")
def sweep_inputs(cuda):
    """The struct-size sweep's inputs at N = 2^26: one 4 GiB source (the
    first 2^26 * k words serve size k), a seeded permutation, both also on
    the host for the oracle."""
    n, kmax = 1 << 26, 16
    d_src = torch.empty(n * kmax, dtype=torch.int32, device=cuda)
    T.fill_words_device(d_src, d_src.numel(), 77)
    perm = T.gen_permutation(n, 4242)
    d_idx = dev(perm)
    d_dst = torch.empty_like(d_src)
    src_host = host(d_src).view(np.uint8)
    yield {"n": n, "d_src": d_src, "d_dst": d_dst, "d_idx": d_idx, "idx_long": d_idx.long(), "perm": perm,
           "src_host": src_host}
    del d_src, d_dst, d_idx
    torch.cuda.empty_cache()


@pytest.mark.parametrize("k", range(1, 17))
def test_full_size_sweep_properties(trove, cuda, sweep_inputs, k):
    # Struct-size sweep at N = 2^26, every k = 1..16: copies are the identity;
    # gathers through the seeded permutation (tile path, coalesced_ptr, the
    # LDS-staged path, and the naive kernel for a few sizes) agree with torch
    # indexing on every row and with the oracle on 4096 sampled rows.
    S = sweep_inputs
    n, s = S["n"], 4 * k
    d_src, d_dst = S["d_src"][:n * k], S["d_dst"][:n * k]
    for variant in ("shfl", "lds"):
        d_dst.zero_()
        T.copy(d_src, d_dst, n, s, variant=variant)
        torch.cuda.synchronize()
        assert _equal_in_chunks(d_dst, d_src), variant
    rows = np.random.default_rng(k).integers(0, n, size=4096)
    want_rows = oracle.gather(S["src_host"][:n * s], S["perm"][rows], s).view(np.int32).reshape(-1, k)
    src2d = d_src.view(n, k)
    step = (1 << 26) // k
    variants = ("tile", "shfl", "lds") + (("direct",) if k in (1, 5, 8, 16) else ())
    for variant in variants:
        d_dst.zero_()
        T.gather(d_src, S["d_idx"], d_dst, n, s, variant=variant)
        torch.cuda.synchronize()
        dst2d = d_dst.view(n, k)
        for r0 in range(0, n, step):
            assert torch.equal(dst2d[r0:r0 + step], src2d[S["idx_long"][r0:r0 + step]]), (variant, r0)
        got_rows = dst2d[torch.from_numpy(rows).long().to(cuda)].cpu().numpy()
        assert np.array_equal(got_rows, want_rows), variant
    if k in (3, 8, 15, 16):
        # The permutation scatter at full size: gathering back through the
        # permutation returns the source.
        ws = torch.empty(n, dtype=torch.int32, device=cuda)
        T.scatter_permutation(d_src, S["d_idx"], d_dst, n, s, workspace=ws)
        torch.cuda.synchronize()
        dst2d = d_dst.view(n, k)
        for r0 in range(0, n, step):
            assert torch.equal(dst2d[S["idx_long"][r0:r0 + step]], src2d[r0:r0 + step]), ("scatter_permutation", r0)
        del ws


@pytest.mark.parametrize("s", [96, 256, 1024])
def test_full_size_large_structs(trove, cuda, s):
    # Large structs at 4 GiB per array: copy is the identity; gathers through a
    # seeded permutation agree with torch indexing; scattering through the
    # permutation and gathering back through it returns the input.
    k = s // 4
    n = (1 << 32) // s
    d_src = torch.empty(n * k, dtype=torch.int32, device=cuda)
    T.fill_words_device(d_src, n * k, 5 + s)
    d_dst = torch.empty_like(d_src)
    T.copy(d_src, d_dst, n, s, variant="shfl")
    torch.cuda.synchronize()
    assert _equal_in_chunks(d_dst, d_src)
    d_idx = dev(T.gen_permutation(n, 99 + s))
    src2d = d_src.view(n, k)
    step = max(1, (1 << 26) // k)
    variants = ("tile", "shfl") + (("lds",) if T.lds_supported(s) else ())
    for variant in variants:
        T.gather(d_src, d_idx, d_dst, n, s, variant=variant)
        torch.cuda.synchronize()
        dst2d = d_dst.view(n, k)
        for r0 in range(0, n, step):
            assert torch.equal(dst2d[r0:r0 + step], src2d[d_idx[r0:r0 + step].long()]), (variant, r0)
    back = torch.empty_like(d_src)
    for variant in variants:
        T.scatter(d_src, d_idx, d_dst, n, s, variant=variant)
        T.gather(d_dst, d_idx, back, n, s, variant=variant)
        torch.cuda.synchronize()
        assert _equal_in_chunks(back, d_src), variant
    del d_src, d_dst, d_idx, back
    torch.cuda.empty_cache()


@pytest.mark.parametrize("s", [4, 64])
def test_gather_scatter_rows_near_the_index_limit(trove, cuda, s):
    # An 8 GiB source: 4-byte rows up to index 2^31 - 2 (the int32 index
    # limit) and 64-byte rows at byte offsets up to 8 GiB, gathered through
    # every path and scattered back, against torch indexing.
    k = s // 4
    n_src = ((1 << 33) // s) - 1
    d_src = torch.empty(n_src * k, dtype=torch.int32, device=cuda)
    T.fill_words_device(d_src, d_src.numel(), 91 + s)
    rng = np.random.default_rng(s)
    n = 64 * 70 + 17
    idx = np.concatenate([n_src - 1 - rng.permutation(n - 5), [0, 1, n_src - 1, n_src // 2, n_src - 2]])
    idx = idx.astype(np.int32)
    d_idx = dev(idx)
    rows = d_src.view(n_src, k)
    want = rows[d_idx.long()]
    for variant in ("tile", "shfl", "lds", "direct"):
        d_dst = torch.zeros(n * k, dtype=torch.int32, device=cuda)
        T.gather(d_src, d_idx, d_dst, n, s, variant=variant)
        torch.cuda.synchronize()
        assert torch.equal(d_dst.view(n, k), want), (s, variant, "gather")
    # Scatter distinct rows near the top back into a copy of their place.
    uniq = np.unique(idx).astype(np.int32)
    m = uniq.size
    d_u = dev(uniq)
    payload = torch.arange(m * k, dtype=torch.int32, device=cuda)
    for variant in ("tile", "shfl", "lds", "direct"):
        T.scatter(payload, d_u, d_src, m, s, variant=variant)
        torch.cuda.synchronize()
        assert torch.equal(rows[d_u.long()], payload.view(m, k)), (s, variant, "scatter")
        T.fill_words_device(payload, payload.numel(), 7 + len(variant))
    del d_src, rows, want
    torch.cuda.empty_cache()


@pytest.mark.parametrize("s", [1, 2, 3, 5, 7, 9, 11, 13, 15])
def test_full_size_block_interface(trove, cuda, s):
    # The block-interface workload at its full shape: 2^28 32-bit elements,
    # s per lane (2^28 % s trailing elements), converged and with one group in
    # eight diverged; the copy must be the identity.
    e = 1 << 28
    d_x = torch.empty(e, dtype=torch.int32, device=cuda)
    T.fill_words_device(d_x, e, 31 + s)
    d_r = torch.empty_like(d_x)
    for divergent in (False, True):
        d_r.fill_(-1)
        T.block_copy(d_x, d_r, e, s, divergent=divergent, seed=77 + s)
        torch.cuda.synchronize()
        assert _equal_in_chunks(d_r, d_x), (s, divergent)
    del d_x, d_r
    torch.cuda.empty_cache()
