"""trove_batch_run_static (the batch entry point that allocates nothing: its
job table lives in the library's static device array) against
trove_batch_create + trove_batch_run(TROVE_BATCH_LAUNCH), bit-exact, on two
calls in a row, the second with fewer and shorter jobs; plus its limits."""
import numpy as np
import pytest
import torch

import trove_amd as T

pytestmark = pytest.mark.gpu


def _jobs(count, seed, max_len):
    """Gather and copy jobs of every struct size and variant, lengths up to
    max_len, on the device: (job tuples into dst_a, job tuples into dst_b)."""
    rng = np.random.default_rng(seed)
    variants = ("tile", "shfl", "lds", "direct", "auto")
    patterns = ("identity", "random", "runs")
    a, b, keep = [], [], []
    for j in range(count):
        s = 4 * (1 + j % 16)
        n = int(rng.integers(1, max_len + 1))
        src = torch.from_numpy(rng.integers(0, 256, size=n * s, dtype=np.uint8)).cuda()
        idx = None
        if j % 4 != 3:  # every fourth job is a copy
            idx = torch.from_numpy(T.gen_indices(n, n, patterns[j % 3], seed * 1000 + j)).cuda()
        dst_a = torch.zeros(n * s, dtype=torch.uint8, device="cuda")
        dst_b = torch.full((n * s,), 0xA5, dtype=torch.uint8, device="cuda")
        v, chunk = variants[(j // 4) % 5], (0, 1, 4, 16)[j % 4]
        a.append((src, idx, dst_a, n, s, v, chunk))
        b.append((src, idx, dst_b, n, s, v, chunk))
        keep.append((src, idx, dst_a, dst_b))
    return a, b, keep


def test_static_batch_matches_allocated_batch_twice(trove, cuda):
    # First call: 300 jobs (five 64-descriptor packs); second call: 37
    # shorter jobs, so the table's tail still holds the first call's
    # descriptors and must not be read.
    for count, max_len, seed in ((300, 5000, 1), (37, 700, 2)):
        ja, jb, keep = _jobs(count, seed, max_len)
        ref = T.Batch(ja)
        ref.run("launch")
        T.batch_run_static(jb)
        torch.cuda.synchronize()
        for j, (_, _, dst_a, dst_b) in enumerate(keep):
            assert torch.equal(dst_a, dst_b), (count, j, ja[j][4], ja[j][5])
        ref.close()


def test_static_batch_limits(trove, cuda):
    src = torch.zeros(4096, dtype=torch.uint8, device=cuda)
    idx = torch.zeros(64, dtype=torch.int32, device=cuda)
    job = (src, idx, src, 64, 32, "tile", 0)
    with pytest.raises(T.TroveError) as e:
        T.batch_run_static([job] * (T.BATCH_STATIC_MAX_JOBS + 1))
    assert e.value.code == T.ERR_FULL
    with pytest.raises(T.TroveError) as e:
        T.batch_run_static([job, (src, idx, src, 64, 68, "tile", 0)])  # a bad job: nothing runs
    assert e.value.code == T.ERR_SIZE
    with pytest.raises(T.TroveError) as e:
        T.batch_run_static([])
    assert e.value.code == T.ERR_INVALID
