"""Writes the golden fixtures of tests/golden/ (run from the repository root).

The reference holds no stored golden files; its tests are known-answer
checks and differential checks. The fixtures restate those known answers as
data, computed here with numpy only (closed forms and numpy indexing), never
with the oracle under test:
  block_kat.npz    tests/block.cu:56-75 -- test_block_write<int,5> writes
                   counting_array(5 * index), so the output is 0, 1, 2, ...;
                   l = 64*3 + 17 keeps the partial last wavefront of the
                   reference's l = 100000*256 + 17.
  counting_kat.npz tests/benchmark.cu:52-88 -- counting structs of k words
                   (utility.h:33-53) are iota words; their per-struct sums
                   (utility.h:55-78) are g*k*k + k*(k-1)/2 mod 2^32.
  index_kat.npz    tests/benchmark.cu:90-128 -- gather r[g] = s[idx[g]] and
                   scatter r[idx[g]] = s[g] on iota structs (fill_test,
                   benchmark.cu:183-189) with a fixed permutation (numpy
                   PCG64, seed 20131201), for k = 1..16 words.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    l, s = 64 * 3 + 17, 5
    np.savez_compressed(os.path.join(HERE, "block_kat.npz"), l=l, s=s,
                        expected=np.arange(l * s, dtype=np.uint32))

    n = 64 * 2 + 17
    counting = {}
    for k in range(1, 17):
        g = np.arange(n, dtype=np.uint64)
        counting[f"words_k{k}"] = np.arange(n * k, dtype=np.uint32)
        counting[f"sums_k{k}"] = ((g * k * k + k * (k - 1) // 2) % (1 << 32)).astype(np.uint32)
    np.savez_compressed(os.path.join(HERE, "counting_kat.npz"), n=n, **counting)

    n = 64 * 3 + 29
    rng = np.random.Generator(np.random.PCG64(20131201))
    perm = rng.permutation(n).astype(np.int32)
    index = {"n": n, "perm": perm}
    for k in range(1, 17):
        src = np.arange(n * k, dtype=np.uint32).reshape(n, k)
        index[f"gather_k{k}"] = src[perm].reshape(-1)
        scat = np.zeros_like(src)
        scat[perm] = src
        index[f"scatter_k{k}"] = scat.reshape(-1)
    np.savez_compressed(os.path.join(HERE, "index_kat.npz"), **index)


if __name__ == "__main__":
    main()
