"""GPU parity of the batched launch and of the gather server.

Batch: every job's output in the one-launch, resident-grid, graph-replayed
and looped modes
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


@pytest.mark.parametrize("mode", ["launch", "resident", "looped"])
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
    for mode in ("launch", "resident", "looped"):
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


def test_batch_graph_first_run_from_two_threads(trove, cuda):
    # The first TROVE_BATCH_GRAPH runs of one plan from two host threads at
    # once: the graph is captured once (under the plan's lock) and both
    # replays, ordered on the default stream, write the right outputs.
    import threading

    jobs = _jobs(64, 12)
    b = _device_jobs(jobs)
    errors = []
    start = threading.Barrier(2)

    def run():
        try:
            start.wait()
            b.run("graph")
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=run) for _ in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    torch.cuda.synchronize()
    _check(jobs, "graph from two threads")
    b.close()


def test_batch_rejects_bad_jobs(trove, cuda):
    src = torch.zeros(1024, dtype=torch.uint8, device=cuda)
    idx = torch.zeros(16, dtype=torch.int32, device=cuda)
    with pytest.raises(T.TroveError) as e:
        T.Batch([(src, idx, src, 16, 68, "tile", 0)])
    assert e.value.code == T.ERR_SIZE
    for variant in (9, T.VARIANT_DIRECT_LAUNCH):  # the launch-equal baseline is not a batch variant
        with pytest.raises(T.TroveError) as e:
            T.Batch([(src, idx, src, 16, 32, variant, 0)])
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
    lat, ids = srv.replay(arrivals, [d.data_ptr() for d in d_reqs], [r.size for r in reqs], on_device=True,
                          release_on_done=False, return_ids=True)
    assert np.all(lat >= 0) and np.all(np.isfinite(lat))
    assert srv.stats()["completed"] == len(reqs)
    assert len(set(ids)) == len(reqs) and 0 not in ids
    for k, r in enumerate(reqs):
        assert np.array_equal(srv.fetch(ids[k], r.size), oracle.gather(host_table, r, rb)), k
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


def test_serve_rejects_device_index_outside_table(trove, cuda):
    # A device-resident request holding indices >= n_rows (and < 0) completes
    # as rejected (ERR_RANGE from poll / output / fetch) while its
    # neighbours in the same steps return the oracle's rows; the bad request
    # spans several steps (slices of 128).
    n_rows, rb = 1 << 14, 32
    host_table, table = _table(n_rows, rb, 21)
    rng = np.random.default_rng(21)
    good = [rng.integers(0, n_rows, size=int(n), dtype=np.int32) for n in (300, 1000, 64, 777)]
    bad = rng.integers(0, n_rows, size=900, dtype=np.int32)
    bad[[5, 640]] = [n_rows, -3]
    d_good = [torch.from_numpy(g).cuda() for g in good]
    d_bad = torch.from_numpy(bad).cuda()
    srv = T.Server(table, n_rows, rb, budget_bytes=1 << 20, max_slots=8, slice_indices=128)
    ids = [srv.submit(d, d.numel(), on_device=True) for d in d_good[:2]]
    bad_id = srv.submit(d_bad, d_bad.numel(), on_device=True)
    ids += [srv.submit(d, d.numel(), on_device=True) for d in d_good[2:]]
    srv.drain()
    for call in (lambda: srv.poll(bad_id), lambda: srv.fetch(bad_id, bad.size), lambda: srv.output(bad_id)):
        with pytest.raises(T.TroveError) as e:
            call()
        assert e.value.code == T.ERR_RANGE
    for i, g in zip(ids, good):
        assert srv.poll(i)[0]
        assert np.array_equal(srv.fetch(i, g.size), oracle.gather(host_table, g, rb))
    srv.release(bad_id)  # its arena rows are returned like any other request's
    # The server keeps serving after a rejection.
    rid = srv.submit(d_good[1], d_good[1].numel(), on_device=True)
    srv.drain()
    assert np.array_equal(srv.fetch(rid, good[1].size), oracle.gather(host_table, good[1], rb))
    srv.close()


def test_serve_copies_host_indices_at_submit(trove, cuda):
    # The caller may overwrite (or drop) its host index array right after
    # submit: the request gathers what was submitted.
    n_rows, rb = 1 << 12, 16
    host_table, table = _table(n_rows, rb, 22)
    srv = T.Server(table, n_rows, rb, budget_bytes=1 << 20, max_slots=4, slice_indices=128)
    idx = np.random.default_rng(22).integers(0, n_rows, size=5000, dtype=np.int32)
    want = oracle.gather(host_table, idx, rb)
    rid = srv.submit(idx, idx.size, on_device=False)
    idx[:] = 0
    del idx
    srv.drain()
    assert np.array_equal(srv.fetch(rid, 5000), want)
    # fetch sizes its buffer from the request's own row count.
    assert srv.rows(rid) == 5000
    assert np.array_equal(srv.fetch(rid), want)
    with pytest.raises(ValueError):
        srv.fetch(rid, 4000)  # a smaller buffer than the request's rows
    srv.release(rid)
    with pytest.raises(T.TroveError) as e:
        srv.rows(rid)
    assert e.value.code == T.ERR_NOTFOUND
    srv.close()


def test_serve_host_indices_beyond_the_staging_pool(trove, cuda):
    # Host indices upload at submission into a staging pool as large as the
    # output budget's rows; submissions that find it full keep a host copy
    # and upload at admission. Both kinds, mixed, gather the right rows, and
    # the caller's arrays are overwritten right after each submission.
    n_rows, rb = 1 << 12, 16
    host_table, table = _table(n_rows, rb, 24)
    srv = T.Server(table, n_rows, rb, budget_bytes=4096 * rb, max_slots=8, slice_indices=512)
    rng = np.random.default_rng(24)
    reqs = [rng.integers(0, n_rows, size=1500, dtype=np.int32) for _ in range(7)]  # 10,500 > 4,096 rows
    wants = [oracle.gather(host_table, r, rb) for r in reqs]
    pending = {}
    for r, want in zip(reqs, wants):
        pending[srv.submit(r, r.size, on_device=False)] = want
        r[:] = 0  # the submission already holds its own indices
    # The outputs of all requests exceed the budget: release each as it
    # finishes so that the queue drains.
    for _ in range(200000):
        if not pending:
            break
        srv.step()
        for rid in list(pending):
            if srv.poll(rid)[0]:
                assert np.array_equal(srv.fetch(rid), pending.pop(rid)), rid
                srv.release(rid)
    assert not pending
    # Once released, the pool takes new submissions again.
    r = rng.integers(0, n_rows, size=3000, dtype=np.int32)
    rid = srv.submit(r, r.size, on_device=False)
    srv.drain()
    assert np.array_equal(srv.fetch(rid), oracle.gather(host_table, r, rb))
    srv.close()


def test_serve_request_ids_stay_unique_across_slot_reuse(trove, cuda):
    # A released request's table slot is reused by the next submission; the
    # old id must then name nothing, and every live id its own request.
    n_rows, rb = 1 << 12, 16
    host_table, table = _table(n_rows, rb, 23)
    srv = T.Server(table, n_rows, rb, budget_bytes=1 << 20, max_slots=4, slice_indices=1024)
    rng = np.random.default_rng(23)
    seen, live = set(), {}
    for round_ in range(6):
        reqs = [rng.integers(0, n_rows, size=int(rng.integers(1, 400)), dtype=np.int32) for _ in range(3)]
        ids = [srv.submit(r, r.size, on_device=False) for r in reqs]
        assert not (set(ids) & seen), round_  # never an id handed out before
        seen.update(ids)
        srv.drain()
        for rid, r in zip(ids, reqs):
            live[rid] = r
        # Release all but the newest request, so later submissions reuse slots.
        for rid in list(live)[:-1]:
            srv.release(rid)
            with pytest.raises(T.TroveError) as e:
                srv.rows(rid)
            assert e.value.code == T.ERR_NOTFOUND
            del live[rid]
        for rid, r in live.items():
            assert np.array_equal(srv.fetch(rid, r.size), oracle.gather(host_table, r, rb)), (round_, rid)
    srv.close()


def test_serve_submit_validates_arrays(trove, cuda):
    n_rows, rb = 1 << 10, 16
    _, table = _table(n_rows, rb, 23)
    srv = T.Server(table, n_rows, rb, budget_bytes=1 << 16, max_slots=4, slice_indices=128)
    with pytest.raises(ValueError):
        srv.submit(np.arange(10, dtype=np.int64), 10, on_device=False)  # int64 would be read as int32 halves
    with pytest.raises(ValueError):
        srv.submit(np.arange(20, dtype=np.int32)[::2], 10, on_device=False)  # not contiguous
    with pytest.raises(ValueError):
        srv.submit(np.arange(5, dtype=np.int32), 10, on_device=False)  # fewer than n
    with pytest.raises(ValueError):
        srv.submit(torch.arange(10, dtype=torch.int64, device=cuda), 10, on_device=True)
    with pytest.raises(ValueError):
        srv.submit(torch.arange(10, dtype=torch.int32), 10, on_device=True)  # host tensor
    with pytest.raises(ValueError):
        srv.submit(torch.arange(5, dtype=torch.int32, device=cuda), 10, on_device=True)
    srv.close()


# ---- the batched and served workloads at their full shapes ----

def test_batched_config_full_shape(trove, cuda):
    # 1,024 jobs, sizes 4..64 B, lengths log-uniform 2^10..2^22, three index
    # patterns, variants and item sizes chosen by the autotuner: every job of
    # the one launch equals torch indexing on the device and the
    # looped launches; 8 jobs equal the oracle; graph replays repeat it.
    from trove_amd import tune
    from trove_amd.workloads import BatchedWorkload

    W = BatchedWorkload(1024, 20260101, torch)
    table = tune.autotune(reps=1, seed=5)
    chosen = set()

    def choose(s, p, L):
        v, c = tune.choose(table, s, p, L)
        chosen.add(v)
        return v, c

    ba = T.Batch(W.jobs(choose, W.dst_a))
    bb = T.Batch(W.jobs(lambda s, p, L: tune.choose(table, s, p, L), W.dst_b))
    ba.run("launch")
    bb.run("looped")
    torch.cuda.synchronize()
    assert len(chosen) >= 2 and W.lengths.min() >= 1 << 10 and W.lengths.max() <= 1 << 22
    idx_long = W.idx.long()
    for j in range(W.n_jobs):
        w0, nw = W.job_words(j)
        k, L = int(W.sizes[j]) // 4, int(W.lengths[j])
        i0 = W.idx_off[j]
        want = W.src[w0:w0 + nw].view(L, k)[idx_long[i0:i0 + L]]
        assert torch.equal(W.dst_a[w0:w0 + nw].view(L, k), want), ("launch", j)
        assert torch.equal(W.dst_b[w0:w0 + nw], W.dst_a[w0:w0 + nw]), ("looped", j)
    for j in list(range(0, W.n_jobs, W.n_jobs // 8))[:8]:
        w0, nw = W.job_words(j)
        want = oracle.gather(W.src[w0:w0 + nw].cpu().numpy().view(np.uint8), W.job_indices(j), int(W.sizes[j]))
        assert np.array_equal(W.dst_a[w0:w0 + nw].cpu().numpy().view(np.uint8), want), ("oracle", j)
    ref = W.dst_a.clone()
    W.dst_b.zero_()
    bb.run("resident")
    torch.cuda.synchronize()
    assert all(torch.equal(ref[i:i + (1 << 26)], W.dst_b[i:i + (1 << 26)])
               for i in range(0, ref.numel(), 1 << 26)), "resident"
    for replay in range(2):
        W.dst_a.zero_()
        torch.cuda.synchronize()
        ba.run("graph")
        torch.cuda.synchronize()
        assert all(torch.equal(ref[i:i + (1 << 26)], W.dst_a[i:i + (1 << 26)])
                   for i in range(0, ref.numel(), 1 << 26)), ("graph", replay)
    ba.close()
    bb.close()
    del W, ref, idx_long
    torch.cuda.empty_cache()


def test_served_config_full_shape(trove, cuda):
    # The 2^24 x 32-B table, a 4 GiB output budget, 64 requests of
    # log-uniform 2^8..2^18 Zipf(1.1) indices on the device, slices of 2^16
    # (so the larger requests span several steps): every request equals the
    # oracle in the mix and alone.
    from trove_amd.workloads import ServedWorkload

    W = ServedWorkload(64, 20260101, torch)
    assert (W.sizes > (1 << 16)).sum() >= 3, "some requests must span steps"
    host_table = W.table.cpu().numpy().view(np.uint8)
    want = [oracle.gather(host_table, W.request(i), W.row_bytes) for i in range(W.n_requests)]
    srv = T.Server(W.table, W.n_rows, W.row_bytes, budget_bytes=4 << 30, max_slots=1024, slice_indices=1 << 16)
    ids = [srv.submit(W.device_request(i), int(W.sizes[i]), on_device=True) for i in range(W.n_requests)]
    srv.drain()
    st = srv.stats()
    assert st["completed"] == W.n_requests and st["steps"] >= 4
    for i, rid in enumerate(ids):
        assert np.array_equal(srv.fetch(rid, int(W.sizes[i])), want[i]), ("mix", i)
        srv.release(rid)
    for i in range(W.n_requests):
        rid = srv.submit(W.device_request(i), int(W.sizes[i]), on_device=True)
        srv.drain()
        assert np.array_equal(srv.fetch(rid, int(W.sizes[i])), want[i]), ("alone", i)
        srv.release(rid)
    srv.close()
    del W, host_table
    torch.cuda.empty_cache()
