"""GPU parity of the batched persistent launch and of the gather server.

Batch: every job's output in the persistent, graph-replayed and looped modes
equals the oracle's gather (or, for copy jobs without indices, copy) of that
job (bit-exact). Graph replays are checked
one by one with eager work and synchronisation between them.
Serve: each request's rows equal the oracle's gather, both when the request
runs alone and when it runs in a mix with others; admission limits (budget,
slots) and index validation behave as documented.
"""
import numpy as np
import pytest
import torch

import oracle
import trove_amd as T

pytestmark = pytest.mark.gpu


def _jobs(count, seed, max_log2=14):
    rng = np.random.default_rng(seed)
    sizes = [4 * (1 + j % 16) for j in range(count)]
    lengths = T.gen_loguniform(count, 1 << 6, 1 << max_log2, seed)
    patterns = ["identity", "random", "runs"]
    variants = ["tile", "shfl", "direct", "lds"]
    jobs = []
    for j in range(count):
        s, n = sizes[j], int(lengths[j])
        src = rng.integers(0, 256, size=n * s, dtype=np.uint8)
        if j % 5 == 4:  # a copy job: no indices
            jobs.append(dict(s=s, n=n, src=src, idx=None, want=oracle.copy(src, n, s),
                             variant=("shfl", "direct", "tile", "lds")[j % 4], chunk=(1, 4, 16, 0)[(j // 20) % 4]))
            continue
        idx = T.gen_indices(n, n, patterns[j % 3], seed * 1000 + j)
        jobs.append(dict(s=s, n=n, src=src, idx=idx, want=oracle.gather(src, idx, s),
                         variant=variants[(j // 3) % 4], chunk=(1, 4, 16, 0)[j % 4]))
    return jobs


def _device_jobs(jobs):
    for J in jobs:
        J["d_src"] = torch.from_numpy(J["src"]).cuda()
        J["d_idx"] = None if J["idx"] is None else torch.from_numpy(J["idx"]).cuda()
        J["d_dst"] = torch.zeros(J["n"] * J["s"], dtype=torch.uint8, device="cuda")
    return T.Batch([(J["d_src"], J["d_idx"], J["d_dst"], J["n"], J["s"], J["variant"], J["chunk"]) for J in jobs])


def _check(jobs, what):
    torch.cuda.synchronize()
    for j, J in enumerate(jobs):
        assert np.array_equal(J["d_dst"].cpu().numpy(), J["want"]), (what, j, J["s"], J["n"], J["variant"])


@pytest.mark.parametrize("mode", ["persistent", "looped"])
def test_batch_modes(trove, cuda, mode):
    jobs = _jobs(96, 7)
    b = _device_jobs(jobs)
    items, nbytes = b.stats()
    assert nbytes == sum((2 * J["s"] + (0 if J["idx"] is None else 4)) * J["n"] for J in jobs) and items > 0
    assert sum(J["idx"] is None for J in jobs) >= 19
    b.run(mode)
    _check(jobs, mode)
    b.close()


def test_batch_tiny_jobs_every_variant(trove, cuda):
    # Jobs shorter than one 64-struct group or a few structs past it: only
    # the per-lane tails run, for gathers and copies of every variant.
    rng = np.random.default_rng(5)
    jobs = []
    for n in (1, 17, 63, 64, 65, 130):
        for s in (4, 12, 32, 60):
            for variant in ("tile", "shfl", "lds", "direct"):
                src = rng.integers(0, 256, size=max(n, 70) * s, dtype=np.uint8)
                idx = rng.integers(0, max(n, 70), size=n, dtype=np.int32)
                jobs.append(dict(s=s, n=n, src=src, idx=idx, want=oracle.gather(src, idx, s), variant=variant,
                                 chunk=1))
                jobs.append(dict(s=s, n=n, src=src, idx=None, want=oracle.copy(src, n, s), variant=variant,
                                 chunk=4))
    b = _device_jobs(jobs)
    for mode in ("persistent", "looped"):
        for J in jobs:
            J["d_dst"].zero_()
        torch.cuda.synchronize()
        b.run(mode)
        _check(jobs, mode)
    b.close()


def test_batch_graph_replays(trove, cuda):
    jobs = _jobs(64, 11)
    b = _device_jobs(jobs)
    for replay in range(3):
        for J in jobs:
            J["d_dst"].zero_()  # eager work between replays
        torch.cuda.synchronize()
        b.run("graph")
        _check(jobs, f"graph replay {replay}")
    b.close()


def test_batch_rejects_bad_jobs(trove, cuda):
    src = torch.zeros(1024, dtype=torch.uint8, device=cuda)
    idx = torch.zeros(16, dtype=torch.int32, device=cuda)
    with pytest.raises(T.TroveError) as e:
        T.Batch([(src, idx, src, 16, 68, "tile", 0)])
    assert e.value.code == T.ERR_SIZE
    with pytest.raises(T.TroveError) as e:
        T.Batch([(src, idx, src, 16, 32, 9, 0)])
    assert e.value.code == T.ERR_VARIANT


def _table(n_rows, row_bytes, seed):
    words = T.gen_words(n_rows * row_bytes // 4, seed)
    return words.view(np.uint8), torch.from_numpy(words.view(np.int32)).cuda()


def test_serve_alone_equals_mix(trove, cuda):
    n_rows, rb = 1 << 16, 32
    host_table, table = _table(n_rows, rb, 5)
    sizes = T.gen_loguniform(40, 1 << 4, 1 << 12, 9)
    reqs = [T.gen_zipf(int(n), n_rows, 1.1, 100 + i) for i, n in enumerate(sizes)]
    want = [oracle.gather(host_table, r, rb) for r in reqs]

    srv = T.Server(table, n_rows, rb, budget_bytes=1 << 20, max_slots=8, slice_indices=256)
    ids = [srv.submit(r, r.size, on_device=False) for r in reqs]
    srv.drain()
    mixed = [srv.fetch(i, r.size) for i, r in zip(ids, reqs)]
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["steps"] > 1
    srv.close()
    for k, r in enumerate(reqs):
        assert np.array_equal(mixed[k], want[k]), ("mix", k)
        alone = T.Server(table, n_rows, rb, budget_bytes=1 << 20, max_slots=8, slice_indices=256)
        rid = alone.submit(r, r.size, on_device=False)
        alone.drain()
        assert np.array_equal(alone.fetch(rid, r.size), mixed[k]), ("alone", k)
        alone.close()


def test_serve_variants_agree(trove, cuda):
    # Every gather path of the server's steps returns the oracle's rows.
    n_rows, rb = 1 << 14, 32
    host_table, table = _table(n_rows, rb, 8)
    sizes = T.gen_loguniform(16, 1 << 4, 1 << 11, 12)
    reqs = [T.gen_zipf(int(n), n_rows, 1.1, 300 + i) for i, n in enumerate(sizes)]
    for variant in ("tile", "shfl", "lds", "direct"):
        srv = T.Server(table, n_rows, rb, budget_bytes=1 << 20, max_slots=4, slice_indices=128, variant=variant)
        ids = [srv.submit(r, r.size, on_device=False) for r in reqs]
        srv.drain()
        for k, (i, r) in enumerate(zip(ids, reqs)):
            assert np.array_equal(srv.fetch(i, r.size), oracle.gather(host_table, r, rb)), (variant, k)
        srv.close()


def test_serve_device_indices_and_replay(trove, cuda):
    n_rows, rb = 1 << 15, 32
    host_table, table = _table(n_rows, rb, 6)
    sizes = T.gen_loguniform(60, 1 << 6, 1 << 13, 10)
    reqs = [T.gen_zipf(int(n), n_rows, 1.1, 200 + i) for i, n in enumerate(sizes)]
    d_reqs = [torch.from_numpy(r).cuda() for r in reqs]
    torch.cuda.synchronize()
    arrivals = np.cumsum(np.random.default_rng(3).exponential(1e-4, size=len(reqs)))
    # The budget holds every output (none is released during the replay).
    assert sum(r.size for r in reqs) * rb <= 1 << 23
    srv = T.Server(table, n_rows, rb, budget_bytes=1 << 23, max_slots=16, slice_indices=1024)
    lat = srv.replay(arrivals, [d.data_ptr() for d in d_reqs], [r.size for r in reqs], on_device=True,
                     release_on_done=False)
    assert np.all(lat >= 0) and np.all(np.isfinite(lat))
    assert srv.stats()["completed"] == len(reqs)
    # Request ids are assigned in submission order starting at 1.
    for k, r in enumerate(reqs):
        assert np.array_equal(srv.fetch(k + 1, r.size), oracle.gather(host_table, r, rb)), k
    srv.close()


def test_serve_replay_reports_a_full_budget(trove, cuda):
    # Kept outputs fill the budget: the replay must stop with ERR_FULL, not spin.
    n_rows, rb = 1 << 12, 16
    _, table = _table(n_rows, rb, 8)
    reqs = [np.arange(300, dtype=np.int32) for _ in range(5)]
    srv = T.Server(table, n_rows, rb, budget_bytes=1000 * rb, max_slots=8, slice_indices=128)
    with pytest.raises(T.TroveError) as e:
        srv.replay(np.zeros(5), [r.ctypes.data for r in reqs], [r.size for r in reqs], on_device=False,
                   release_on_done=False)
    assert e.value.code == T.ERR_FULL
    srv.close()


def test_serve_admission_limits(trove, cuda):
    n_rows, rb = 4096, 16
    _, table = _table(n_rows, rb, 7)
    srv = T.Server(table, n_rows, rb, budget_bytes=1000 * rb, max_slots=4, slice_indices=128)
    big = np.zeros(1001, dtype=np.int32)
    with pytest.raises(T.TroveError) as e:
        srv.submit(big, big.size, on_device=False)
    assert e.value.code == T.ERR_BUDGET
    bad = np.array([0, 4096], dtype=np.int32)
    with pytest.raises(T.TroveError) as e:
        srv.submit(bad, bad.size, on_device=False)
    assert e.value.code == T.ERR_RANGE
    a = np.arange(600, dtype=np.int32)
    b = np.arange(600, dtype=np.int32)[::-1].copy()
    ra = srv.submit(a, a.size, on_device=False)
    rb_ = srv.submit(b, b.size, on_device=False)
    # 600 + 600 rows exceed the 1000-row budget: b waits until a is released.
    with pytest.raises(T.TroveError) as e:
        srv.drain()
    assert e.value.code == T.ERR_FULL
    assert srv.poll(ra)[0] and not srv.poll(rb_)[0]
    with pytest.raises(T.TroveError) as e:
        srv.release(rb_)
    assert e.value.code == T.ERR_AGAIN
    srv.release(ra)
    srv.drain()
    assert srv.poll(rb_)[0]
    e0 = srv.submit(np.zeros(0, dtype=np.int32), 0, on_device=False)
    assert srv.poll(e0)[0]
    srv.close()
