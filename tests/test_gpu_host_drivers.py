"""The C++ host drivers (tests/host/trove_benchmark.cpp, trove_block.cpp),
which run the reference's benchmark and block test programs
(tests/benchmark.cu, tests/block.cu) through the C ABI: every benchmark line
must read "Results passed" (the Trove kernel's output equals the naive
kernel's) and both block lines "pass: true"."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HOST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host")
NAMES = ["Contiguous SHFL Store", "Contiguous Direct Store", "Contiguous SHFL Load", "Contiguous Direct Load",
         "SHFL Scatter", "Direct Scatter", "SHFL Gather", "Direct Gather", "Tile Scatter", "Tile Gather",
         "LDS Scatter", "LDS Gather"]


def _run(args, timeout=240):
    exe = os.path.join(HOST, args[0])
    assert os.path.exists(exe), f"{exe} not built (run __graft_entry__.build())"
    return subprocess.run([exe, *args[1:]], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("tile", [None, "32", "16", "2"])
def test_benchmark_driver_results_pass(cuda, tile):
    # 2^20 + 37 structs: the last wavefront of every kernel is partial. tile:
    # the reference program's TILE_SIZE (32 by default there, tests/Makefile:4)
    # for the contiguous store and load benchmarks; None = 64.
    args = ["trove_benchmark", "--n", str((1 << 20) + 37), "--iterations", "2"]
    names = NAMES
    if tile is not None:
        # The tile width changes only the contiguous benchmarks: run those.
        args += ["--tile", tile, "--contiguous-only", "1"]
        names = NAMES[:4]
    p = _run(args)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    lines = [ln.split(", ") for ln in p.stdout.strip().splitlines()]
    assert [(ln[0], int(ln[1])) for ln in lines] == [(n, k) for n in names for k in range(1, 17)]
    assert all(ln[3] == "Results passed" and float(ln[2]) > 0 for ln in lines), p.stdout


@pytest.mark.parametrize("args", [[], ["--s", "7", "--divergent", "1"], ["--s", "1", "--l", "4097"]])
def test_block_driver_passes(cuda, args):
    p = _run(["trove_block", *args])
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.splitlines() == ["test_block_write results pass: true", "test_block_copy results pass: true"]
