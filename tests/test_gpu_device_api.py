"""GPU unit checks of the public device API (tests/host/devtests.hip).

* r2c_compute_indices + r2c_warp_transpose (explicit indices) map the
  column-side layout to the row-side one and c2r_* restores it, for tile
  widths 64..2, odd / power-of-two / composite m and 4-, 8- and 16-byte
  elements (closed-form post-condition, trove/transpose.h:629-696);
* rotate with a per-lane amount equals np.roll (rotate.h:111-115);
* warp_converged<N> under arbitrary EXEC masks equals the truth table of
  "every lane of my aligned N-group is active" (warp.h:56-93);
* thread_tile<N>::shfl reads from the addressed rank of the caller's tile;
* trove::load / trove::store under arbitrary EXEC masks move exactly the
  active lanes' structs (the tile cascade, aos.h:258-290);
* load_warp_contiguous_lds / store_warp_contiguous_lds (trove/lds.h, no
  reference counterpart) put struct l in lane l and write it back;
* warp_load / warp_store / uncoalesced_store (memory.h:136-168) move the
  words the oracle's restatement moves, for the coalesced pattern and for
  arbitrary per-lane offsets and strides, plain and volatile overloads;
* static_mod_inverse (with and without a start value), static_gcd,
  static_log, is_power_of_two, is_odd and inverse (static_mod_inverse.h:
  32-52, static_gcd.h, utility.h:80-126), evaluated by the product headers,
  equal the oracle's tables.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "libtrove_devtests.so")


@pytest.fixture(scope="module")
def dt(cuda):
    assert os.path.exists(LIB), f"{LIB} not built (run __graft_entry__.build())"
    L = ctypes.CDLL(LIB)
    vp, i, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
    L.devtest_transpose.argtypes = [i, i, i, vp, vp]
    L.devtest_transpose_self.argtypes = [i, i, i, vp, vp]
    L.devtest_transpose_self.restype = i
    L.devtest_rotate.argtypes = [i, vp]
    L.devtest_converged.argtypes = [u64, vp]
    L.devtest_shfl.argtypes = [i, vp, vp]
    L.devtest_cascade.argtypes = [i, u64, i, vp, vp, vp]
    L.devtest_cascade_policy.argtypes = [i, u64, i, i, vp, vp, vp]
    L.devtest_cascade_policy.restype = i
    L.devtest_inplace.argtypes = [i, i, vp, vp]
    L.devtest_inplace.restype = i
    L.devtest_shared_gather.argtypes = [i, u64, vp, vp, vp]
    L.devtest_shared_gather.restype = i
    L.devtest_mixed.argtypes = [i, u64, vp, vp, vp, vp]
    L.devtest_mixed.restype = i
    L.devtest_lds.argtypes = [i, i, i, vp, vp, vp]
    L.devtest_lds.restype = i
    L.devtest_user_struct.argtypes = [vp, vp, vp, vp, vp]
    L.devtest_user_struct.restype = i
    L.devtest_memory.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp]
    L.devtest_memory.restype = i
    L.devtest_static_table_sizes.argtypes = [vp]
    L.devtest_static_table_sizes.restype = None
    L.devtest_static_tables.argtypes = [vp]
    L.devtest_static_tables.restype = i
    for f in (L.devtest_transpose, L.devtest_rotate, L.devtest_converged, L.devtest_shfl, L.devtest_cascade):
        f.restype = i
    return L


TRANSPOSE_CASES = (
    [(n, m, 4) for n in (64, 32, 16) for m in (2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16) if m <= n]
    + [(8, m, 4) for m in (2, 3, 5, 6, 8)] + [(4, 2, 4), (4, 3, 4), (4, 4, 4), (2, 2, 4)]
    + [(64, m, 8) for m in (2, 3, 6, 12)] + [(64, m, 16) for m in (2, 5, 6, 16)] + [(32, 6, 16)]
    + [(64, 24, 4), (64, 33, 4), (64, 64, 4)]
    + [(64, 24, 16), (64, 32, 16), (32, 32, 16), (64, 48, 16), (64, 64, 16)]
)


@pytest.mark.parametrize("n,m,wb", TRANSPOSE_CASES)
def test_public_transpose_api(dt, n, m, wb):
    r2c = torch.zeros(64 * m, dtype=torch.int32, device="cuda")
    c2r = torch.zeros(64 * m, dtype=torch.int32, device="cuda")
    assert dt.devtest_transpose(n, m, wb, r2c.data_ptr(), c2r.data_ptr()) == 0
    rank = np.arange(64) % n
    want_row = (rank[:, None] * m + np.arange(m)[None, :]).astype(np.int64)
    want_col = (rank[:, None] + np.arange(m)[None, :] * n).astype(np.int64)
    got_r2c = r2c.cpu().numpy().view(np.uint32).reshape(64, m).astype(np.int64)
    got_c2r = c2r.cpu().numpy().view(np.uint32).reshape(64, m).astype(np.int64)
    assert np.array_equal(got_r2c, want_row), (n, m, wb)
    assert np.array_equal(got_c2r, want_col), (n, m, wb)


SELF_CASES = [(64, 64, 4), (64, 64, 16), (32, 32, 4), (32, 32, 16), (16, 16, 4), (16, 16, 8), (8, 8, 4), (4, 4, 4),
              (2, 2, 4), (2, 2, 16), (64, 5, 4), (64, 6, 16), (32, 12, 4), (64, 48, 16)]


@pytest.mark.parametrize("n,m,wb", SELF_CASES)
def test_self_contained_transpose(dt, n, m, wb):
    # Square tiles run the butterfly transpose (v_permlane32/16_swap + DPP),
    # the others the partner schedule; both must meet the same post-condition.
    r2c = torch.zeros(64 * m, dtype=torch.int32, device="cuda")
    c2r = torch.zeros(64 * m, dtype=torch.int32, device="cuda")
    assert dt.devtest_transpose_self(n, m, wb, r2c.data_ptr(), c2r.data_ptr()) == 0
    rank = np.arange(64) % n
    want_row = (rank[:, None] * m + np.arange(m)[None, :]).astype(np.int64)
    want_col = (rank[:, None] + np.arange(m)[None, :] * n).astype(np.int64)
    got_r2c = r2c.cpu().numpy().view(np.uint32).reshape(64, m).astype(np.int64)
    got_c2r = c2r.cpu().numpy().view(np.uint32).reshape(64, m).astype(np.int64)
    assert np.array_equal(got_r2c, want_row), (n, m, wb)
    assert np.array_equal(got_c2r, want_col), (n, m, wb)


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 13, 15, 16])
def test_rotate_per_lane(dt, m):
    out = torch.zeros(64 * m, dtype=torch.int32, device="cuda")
    assert dt.devtest_rotate(m, out.data_ptr()) == 0
    got = out.cpu().numpy().reshape(64, m)
    lim = 1
    while lim < m:
        lim <<= 1
    for lane in range(64):
        t = 100 * lane + np.arange(m)
        assert np.array_equal(got[lane], np.roll(t, -((lane % lim) % m))), (m, lane)


def _masks():
    rng = np.random.default_rng(5)
    fixed = [0xFFFFFFFFFFFFFFFF, 0x00000000FFFFFFFF, 0xFFFFFFFF00000000, 0x0F0F0F0F0F0F0F0F,
             0x5555555555555555, 0x1, 0x3, 0xFFFF0000FFFF0000, 0x00FF00FF00FF00FF, 0x8000000000000001]
    return fixed + [int(x) for x in rng.integers(0, 2**63, size=20, dtype=np.int64)] + \
        [(1 << 64) - 1 - (1 << int(b)) for b in rng.integers(0, 64, size=5)]


@pytest.mark.parametrize("mask", _masks())
def test_warp_converged_truth_table(dt, mask):
    out = torch.zeros(64 * 6, dtype=torch.int32, device="cuda")
    assert dt.devtest_converged(mask, out.data_ptr()) == 0
    got = out.cpu().numpy().reshape(64, 6)
    active = [(mask >> l) & 1 for l in range(64)]
    for lane in range(64):
        if not active[lane]:
            assert (got[lane] == -1).all()
            continue
        for k, N in enumerate((64, 32, 16, 8, 4, 2)):
            base = lane & ~(N - 1)
            want = int(all(active[base:base + N]))
            assert got[lane, k] == want, (hex(mask), lane, N)


@pytest.mark.parametrize("n", [64, 32, 16, 8, 4, 2])
def test_tile_shfl(dt, n):
    src = np.random.default_rng(n).integers(0, n, size=64).astype(np.int32)
    d_src = torch.from_numpy(src).cuda()
    out = torch.zeros(128, dtype=torch.int32, device="cuda")
    assert dt.devtest_shfl(n, d_src.data_ptr(), out.data_ptr()) == 0
    got = out.cpu().numpy().view(np.uint32).reshape(64, 2)
    for lane in range(64):
        from_lane = (lane & ~(n - 1)) + int(src[lane])
        v = from_lane * 0x100000001 + 7
        assert got[lane, 0] == v & 0xFFFFFFFF and got[lane, 1] == v >> 32, (n, lane)


@pytest.mark.parametrize("s", [8, 12, 20, 32, 48, 60, 96, 384, 512, 1024])
def test_cascade_under_masks(dt, s):
    rng = np.random.default_rng(s)
    k = s // 4
    for mask in _masks()[:18]:
        for scatter in (0, 1):
            for perm in (np.arange(64), rng.permutation(64)):
                perm = perm.astype(np.int32)
                src = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32)
                init = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32)
                d_src = torch.from_numpy(src.view(np.int32)).cuda()
                d_dst = torch.from_numpy(init.view(np.int32)).cuda()
                d_perm = torch.from_numpy(perm).cuda()
                assert dt.devtest_cascade(s, mask, scatter, d_src.data_ptr(), d_perm.data_ptr(),
                                          d_dst.data_ptr()) == 0
                got = d_dst.cpu().numpy().view(np.uint32).reshape(64, k)
                want = init.reshape(64, k).copy()
                src2 = src.reshape(64, k)
                for lane in range(64):
                    if (mask >> lane) & 1:
                        if scatter:
                            want[perm[lane]] = src2[lane]
                        else:
                            want[lane] = src2[perm[lane]]
                assert np.array_equal(got, want), (s, hex(mask), scatter)


@pytest.mark.parametrize("s", [8, 20, 32, 60, 64, 96, 256])
def test_ref_assignment_in_place(dt, s):
    # coalesced_ref = coalesced_ref inside one array: a wavefront permutes 64
    # structs in place; every lane reads the structs as they were before.
    rng = np.random.default_rng(950 + s)
    k = s // 4
    for trial in range(4):
        for scatter in (0, 1):
            perm = rng.permutation(64).astype(np.int32)
            old = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32).reshape(64, k)
            d_buf = torch.from_numpy(old.reshape(-1).view(np.int32).copy()).cuda()
            assert dt.devtest_inplace(s, scatter, d_buf.data_ptr(), torch.from_numpy(perm).cuda().data_ptr()) == 0
            got = d_buf.cpu().numpy().view(np.uint32).reshape(64, k)
            want = np.empty_like(old)
            if scatter:
                want[perm] = old
            else:
                want[:] = old[perm]
            assert np.array_equal(got, want), (s, trial, scatter)


@pytest.mark.parametrize("s", [8, 20, 32, 60, 64, 96, 256])
def test_streaming_policy_cascade(dt, s):
    # The cascade with nontemporal accesses (trove::streaming) moves the same
    # structs as the cached one: trove::load / store<T, streaming>, two
    # streaming coalesced_ptrs, a cached source assigned to a streaming
    # destination, and two cached coalesced_ptrs; gathers and scatters under
    # arbitrary lane masks.
    rng = np.random.default_rng(900 + s)
    k = s // 4
    for mask in _masks()[:10]:
        for scatter in (0, 1):
            for mode in (0, 1, 2, 3):
                perm = rng.permutation(64).astype(np.int32)
                src = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32)
                init = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32)
                d_dst = torch.from_numpy(init.view(np.int32)).cuda()
                assert dt.devtest_cascade_policy(s, mask, scatter, mode, torch.from_numpy(src.view(np.int32)).cuda().data_ptr(),
                                                 torch.from_numpy(perm).cuda().data_ptr(), d_dst.data_ptr()) == 0
                got = d_dst.cpu().numpy().view(np.uint32).reshape(64, k)
                want = init.reshape(64, k).copy()
                src2 = src.reshape(64, k)
                for lane in range(64):
                    if (mask >> lane) & 1:
                        if scatter:
                            want[perm[lane]] = src2[lane]
                        else:
                            want[lane] = src2[perm[lane]]
                assert np.array_equal(got, want), (s, hex(mask), scatter, mode)


@pytest.mark.parametrize("s", [8, 20, 32, 60, 64])
def test_coalesced_ptr_over_lds(dt, s):
    # The dynamic path on pointers into a __shared__ array (LDS aperture).
    rng = np.random.default_rng(100 + s)
    k = s // 4
    for mask in _masks()[:12]:
        perm = rng.permutation(64).astype(np.int32)
        src = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32)
        init = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32)
        d_dst = torch.from_numpy(init.view(np.int32)).cuda()
        assert dt.devtest_shared_gather(s, mask, torch.from_numpy(src.view(np.int32)).cuda().data_ptr(),
                                        torch.from_numpy(perm).cuda().data_ptr(), d_dst.data_ptr()) == 0
        got = d_dst.cpu().numpy().view(np.uint32).reshape(64, k)
        want = init.reshape(64, k).copy()
        for lane in range(64):
            if (mask >> lane) & 1:
                want[lane] = src.reshape(64, k)[perm[lane]]
        assert np.array_equal(got, want), (s, hex(mask))


@pytest.mark.parametrize("s", [4, 8, 12, 20, 32, 60, 64, 96, 256])
def test_lds_staged_api(dt, s):
    # load_warp_contiguous_lds gives lane l struct l of its tile (checked by a
    # plain per-lane store) and store_warp_contiguous_lds writes it back, with
    # several wavefronts per work-group on private stages; tile bases 16-byte
    # aligned (dwordx4 chunks) and 4 or 8 bytes off (dword chunks). Bytes
    # around the tiles stay untouched.
    k = s // 4
    tiles = 8
    rng = np.random.default_rng(500 + s)
    for off in (0, 4, 8):
        nbytes = off + tiles * 64 * s + 16
        src = rng.integers(0, 256, size=nbytes, dtype=np.uint8)
        init = rng.integers(0, 256, size=nbytes, dtype=np.uint8)
        d_src = torch.from_numpy(src).cuda()
        d_dst = torch.from_numpy(init.copy()).cuda()
        d_got = torch.zeros(tiles * 64 * k, dtype=torch.int32, device="cuda")
        assert dt.devtest_lds(s, off, tiles, d_src.data_ptr(), d_got.data_ptr(), d_dst.data_ptr()) == 0
        body = src[off:off + tiles * 64 * s]
        assert np.array_equal(d_got.cpu().numpy().view(np.uint8), body), (s, off, "lane l holds struct l")
        want = init.copy()
        want[off:off + tiles * 64 * s] = body
        assert np.array_equal(d_dst.cpu().numpy(), want), (s, off, "store")


def test_user_struct_of_mixed_fields(dt):
    # struct {float x, y, z; int id; double mass;} (24 bytes, 8-byte words):
    # coalesced_ptr gather (ref = ref), load/store_warp_contiguous and the
    # LDS-staged pair move it byte for byte, NaN and denormal floats included.
    rng = np.random.default_rng(77)
    src = rng.integers(0, 256, size=64 * 24, dtype=np.uint8)
    f = src.view(np.uint32).reshape(64, 6)
    f[::5, 0] = 0x7FC00001  # quiet NaN with payload
    f[1::7, 1] = 0x00000003  # denormal
    f[2::9, 2] = 0xFF800000  # -inf
    perm = rng.permutation(64).astype(np.int32)
    d_src = torch.from_numpy(src).cuda()
    outs = [torch.zeros(64 * 24, dtype=torch.uint8, device="cuda") for _ in range(3)]
    assert dt.devtest_user_struct(d_src.data_ptr(), torch.from_numpy(perm).cuda().data_ptr(),
                                  *(o.data_ptr() for o in outs)) == 0
    rows = src.reshape(64, 24)
    assert np.array_equal(outs[0].cpu().numpy().reshape(64, 24), rows[perm]), "coalesced_ptr gather"
    assert np.array_equal(outs[1].cpu().numpy(), src), "warp-contiguous copy"
    assert np.array_equal(outs[2].cpu().numpy(), src), "LDS-staged copy"


@pytest.mark.parametrize("s", [12, 20, 32, 48, 96])
def test_dynamic_path_pointers_a_word_apart(dt, s):
    # Lanes of one tile point into two views of a buffer that start one word
    # (not a whole struct) apart: the struct-offset shortcuts must not apply.
    word = 16 if s % 16 == 0 else 8 if s % 8 == 0 else 4
    rng = np.random.default_rng(300 + s)
    k = s // 4
    buf = rng.integers(0, 256, size=64 * s + 64, dtype=np.uint8)
    d_buf = torch.from_numpy(buf).cuda()
    for mask in _masks()[:10]:
        for pattern in ("alternate", "halves", "random"):
            perm = rng.permutation(64).astype(np.int32)
            if pattern == "alternate":
                shift = (np.arange(64) % 2) * word
            elif pattern == "halves":
                shift = (np.arange(64) >= 32) * word
            else:
                shift = rng.integers(0, 2, size=64) * word
            shift = shift.astype(np.int32)
            init = rng.integers(0, 2**32, size=64 * k, dtype=np.uint64).astype(np.uint32)
            d_dst = torch.from_numpy(init.view(np.int32)).cuda()
            assert dt.devtest_mixed(s, mask, d_buf.data_ptr(), torch.from_numpy(perm).cuda().data_ptr(),
                                    torch.from_numpy(shift).cuda().data_ptr(), d_dst.data_ptr()) == 0
            got = d_dst.cpu().numpy().view(np.uint32).reshape(64, k)
            want = init.reshape(64, k).copy()
            for lane in range(64):
                if (mask >> lane) & 1:
                    off = int(shift[lane]) + int(perm[lane]) * s
                    want[lane] = buf[off:off + s].view(np.uint32)
            assert np.array_equal(got, want), (s, hex(mask), pattern)


def _memory_case(rng, s, lanes, coalesced):
    # Distinct positions across lanes: each lane owns a disjoint segment and
    # walks it with its own stride >= 1 (or offset = lane, stride = 64).
    if coalesced:
        return np.arange(lanes, dtype=np.int32), np.full(lanes, lanes, dtype=np.int32), lanes * s
    seg = 4 * s + 3
    stride = rng.integers(1, 5, size=lanes)
    lead = np.array([rng.integers(0, seg - (s - 1) * int(st)) for st in stride])
    return ((rng.permutation(lanes) * seg + lead).astype(np.int32), stride.astype(np.int32), lanes * seg)


@pytest.mark.parametrize("eb,s,vol", [(4, 1, 0), (4, 2, 0), (4, 5, 0), (4, 16, 0), (8, 3, 0), (16, 2, 0),
                                      (16, 5, 0), (4, 5, 1), (8, 3, 1)])
def test_warp_memory_api(dt, eb, s, vol):
    import oracle

    rng = np.random.default_rng(eb * 1000 + s * 10 + vol)
    for coalesced in (True, False):
        offset, stride, span = _memory_case(rng, s, 64, coalesced)
        src = rng.integers(0, 256, size=span * eb, dtype=np.uint8)
        init = rng.integers(0, 256, size=span * eb, dtype=np.uint8)
        d_src, d_stored = torch.from_numpy(src).cuda(), torch.from_numpy(init.copy()).cuda()
        d_loaded = torch.zeros(64 * s * eb, dtype=torch.uint8, device="cuda")
        d_unc = torch.zeros(64 * s * eb, dtype=torch.uint8, device="cuda")
        assert dt.devtest_memory(eb, s, vol, d_src.data_ptr(), torch.from_numpy(offset).cuda().data_ptr(),
                                 torch.from_numpy(stride).cuda().data_ptr(), d_loaded.data_ptr(),
                                 d_stored.data_ptr(), d_unc.data_ptr()) == 0
        loaded, stored, unc = oracle.warp_memory(src, eb, s, offset, stride, init)
        assert np.array_equal(d_loaded.cpu().numpy().reshape(64, -1), loaded), (eb, s, vol, coalesced, "load")
        assert np.array_equal(d_stored.cpu().numpy(), stored), (eb, s, vol, coalesced, "warp_store")
        assert np.array_equal(d_unc.cpu().numpy().reshape(64, -1), unc), (eb, s, vol, coalesced, "uncoalesced")


def test_static_tables(dt):
    import oracle

    L = oracle.lib()
    sizes = np.zeros(6, dtype=np.int32)
    dt.devtest_static_table_sizes(sizes.ctypes.data)
    out = torch.zeros(int(sizes.sum()), dtype=torch.int32, device="cuda")
    assert dt.devtest_static_tables(out.data_ptr()) == 0
    parts = np.split(out.cpu().numpy(), np.cumsum(sizes)[:-1])
    inv, inv_r, gcd, log, bits, inverse = parts
    inv = inv.reshape(65, 64)
    for a in range(65):
        for m in range(1, 65):
            assert inv[a, m - 1] == L.oracle_mod_inverse(a, m), ("static_mod_inverse", a, m)
    inv_r = inv_r.reshape(6, 33, 32)
    for r in range(1, 7):
        for a in range(33):
            for m in range(1, 33):
                assert inv_r[r - 1, a, m - 1] == L.oracle_mod_inverse_from(a, m, r), ("start", a, m, r)
    gcd = gcd.reshape(65, 65)
    for u in range(1, 65):
        for v in range(1, 65):
            assert gcd[u, v] == L.oracle_gcd(u, v), ("static_gcd", u, v)
    # gcd(0, v) = v and gcd(u, 0) = u (the reference's zero cases are bool-typed).
    assert np.array_equal(gcd[0, 1:], np.arange(1, 65)) and np.array_equal(gcd[1:, 0], np.arange(1, 65))
    for m in range(1, 1025):
        assert log[m - 1] == L.oracle_floor_log2(m), ("static_log", m)
    for m in range(0, 1025):
        assert bits[m] == L.oracle_pow2_odd_bits(m), ("is_power_of_two / is_odd", m)
    fn = [(5 * p + 3) % 64 for p in range(200)]
    for x in range(64):
        assert inverse[x] == oracle.inverse(fn, x, 0) and inverse[64 + x] == oracle.inverse(fn, x, 10), x
