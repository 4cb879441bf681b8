"""CPU-only checks of the product: the C ABI library loads and exports what
include/trove_amd.h declares, argument validation happens before any GPU
call, the input generators are deterministic, the device library's schedule
meets the transposition post-conditions on the host, and kernels written
against the reference's documented API compile unchanged for gfx950."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import trove_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def test_library_exports_every_declared_function():
    L = T.lib()
    names = T.declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(T._SIGNATURES) >= set(names), set(names) - set(T._SIGNATURES)


def test_error_strings_and_version():
    assert T.strerror(T.OK) == "success"
    for code in range(-11, 0):
        assert T.strerror(code) != "unknown error"
    assert T.strerror(-99) == "unknown error"
    assert T.lib().trove_version() >= 10000


def test_struct_size_table():
    L = T.lib()
    for b in T.STRUCT_SIZES:
        assert L.trove_struct_size_supported(b) == 1
    for b in (0, 1, 3, 68, 100, 320, 1028, 2048):
        assert L.trove_struct_size_supported(b) == 0
    assert [L.trove_struct_word_bytes(b) for b in (4, 8, 12, 16, 20, 24, 32, 6, 2, 384, 1024)] == \
        [4, 8, 4, 16, 4, 8, 16, 1, 1, 16, 16]


def test_size_shards_match_build():
    """build.py compiles one object per size shard that csrc/common.h defines."""
    import re
    from trove_amd import build
    text = open(os.path.join(build.CSRC, "common.h")).read()
    shards = int(re.search(r"#define TROVE_SIZE_SHARDS (\d+)", text).group(1))
    assert shards == build.SIZE_SHARDS
    sizes = sorted(int(x) for k in range(shards)
                   for x in re.findall(r"X\((\d+)\)", re.search(rf"#define TROVE_SIZES_SHARD{k}\(X\)(.*)", text).group(1)))
    assert sizes == sorted(T.STRUCT_SIZES)


def test_bench_kernels_limited_to_256_bytes():
    L = T.lib()
    fake = 0x10000  # never dereferenced: the size check returns before any launch
    for b in (384, 512, 768, 1024):
        assert L.trove_store_counting(fake, 100, b, 1, 0, None) == T.ERR_SIZE
        assert L.trove_load_sum(fake, fake, 100, b, 1, 0, None) == T.ERR_SIZE


def test_lds_variant_limits():
    # The LDS-staged variant: whole wavefronts only (tile 0 or 64), struct
    # sizes that are multiples of 4 up to 512 bytes, and not for the store /
    # load benchmarks. Every call below returns before any launch.
    L = T.lib()
    fake = 0x10000
    assert L.trove_copy(fake, fake, 10, 20, T.VARIANT_LDS, 32, None) == T.ERR_VARIANT
    for b in (768, 1024, 2, 6):
        assert not T.lds_supported(b)
        assert L.trove_copy(fake, fake, 10, b, T.VARIANT_LDS, 64, None) == T.ERR_VARIANT
        assert L.trove_gather(fake, fake, fake, 10, b, T.VARIANT_LDS, 0, None) == T.ERR_VARIANT
        assert L.trove_scatter(fake, fake, fake, 10, b, T.VARIANT_LDS, 0, None) == T.ERR_VARIANT
    assert L.trove_store_counting(fake, 10, 20, T.VARIANT_LDS, 0, None) == T.ERR_VARIANT
    assert L.trove_load_sum(fake, fake, 10, 20, T.VARIANT_LDS, 0, None) == T.ERR_VARIANT
    assert [b for b in T.STRUCT_SIZES if T.lds_supported(b)] == \
        [b for b in T.STRUCT_SIZES if b % 4 == 0 and b <= 512]
    assert all(T.word_bytes(b) == L.trove_struct_word_bytes(b) for b in T.STRUCT_SIZES)


def test_auto_variant_table():
    # What AUTO resolves to, per op and size (no GPU needed).
    assert [T.auto_variant("copy", b) for b in (4, 8, 12, 16, 20, 32, 60, 96, 128, 160, 256, 1024, 6)] == \
        ["lds", "lds", "lds", "lds", "shfl", "lds", "shfl", "lds", "lds", "lds", "shfl", "shfl", "shfl"]
    assert [T.auto_variant("gather", b) for b in (4, 16, 20, 48, 64, 192, 224, 384, 512, 2)] == \
        ["tile", "lds", "tile", "lds", "lds", "lds", "shfl", "lds", "tile", "direct"]
    assert [T.auto_variant("scatter", b) for b in (4, 8, 16, 20, 64, 6)] == \
        ["direct", "lds", "lds", "tile", "shfl", "direct"]
    assert T.lib().trove_auto_variant(3, 20) == T.ERR_INVALID
    assert T.lib().trove_auto_variant(0, 68) == T.ERR_SIZE


def test_validation_before_launch():
    L = T.lib()
    fake = 0x10000  # never dereferenced: every call below fails validation
    assert L.trove_copy(fake, fake, 10, 68, 1, 64, None) == T.ERR_SIZE
    assert L.trove_copy(fake, fake, -1, 20, 1, 64, None) == T.ERR_INVALID
    assert L.trove_copy(fake, fake, 10, 20, 7, 64, None) == T.ERR_VARIANT
    assert L.trove_copy(fake, fake, 10, 20, T.VARIANT_TILE, 64, None) == T.ERR_VARIANT
    assert L.trove_copy(fake, fake, 10, 20, 1, 48, None) == T.ERR_VARIANT
    assert L.trove_copy(fake + 2, fake, 10, 20, 1, 64, None) == T.ERR_ALIGN
    assert L.trove_copy(fake + 4, fake, 10, 32, 1, 64, None) == T.ERR_ALIGN
    assert L.trove_copy(0, fake, 10, 20, 1, 64, None) == T.ERR_INVALID
    assert L.trove_copy(0, 0, 0, 20, 1, 64, None) == T.OK  # nothing to do
    assert L.trove_gather(fake, fake + 2, fake, 10, 20, 3, 0, None) == T.ERR_ALIGN
    assert L.trove_gather(fake, 0, fake, 10, 20, 3, 0, None) == T.ERR_INVALID
    assert L.trove_scatter(fake, fake, fake, 10, 20, 9, 0, None) == T.ERR_VARIANT
    assert L.trove_gather(fake, fake, fake, 10, 20, T.VARIANT_DIRECT_LAUNCH + 1, 0, None) == T.ERR_VARIANT
    # The launch-equal naive variant is for copy, gather and scatter only.
    assert L.trove_store_counting(fake, 10, 20, T.VARIANT_DIRECT_LAUNCH, 0, None) == T.ERR_VARIANT
    assert L.trove_load_sum(fake, fake, 10, 20, T.VARIANT_DIRECT_LAUNCH, 0, None) == T.ERR_VARIANT
    assert L.trove_gather(0, 0, 0, 0, 20, 3, 0, None) == T.OK
    assert L.trove_store_counting(fake, 10, 6, 1, 0, None) == T.ERR_SIZE
    assert L.trove_load_sum(fake, fake + 1, 10, 20, 1, 0, None) == T.ERR_ALIGN
    for bad_tile in (1, 3, 48, 128, -64):  # gather / scatter tile widths
        assert L.trove_gather(fake, fake, fake, 10, 20, 1, bad_tile, None) == T.ERR_VARIANT
        assert L.trove_scatter(fake, fake, fake, 10, 20, 1, bad_tile, None) == T.ERR_VARIANT
    # Sub-wavefront tiles are coalesced_ptr's: not for the tile or LDS-staged paths.
    for v in (T.VARIANT_TILE, T.VARIANT_LDS):
        assert L.trove_gather(fake, fake, fake, 10, 20, v, 32, None) == T.ERR_VARIANT
        assert L.trove_scatter(fake, fake, fake, 10, 20, v, 16, None) == T.ERR_VARIANT
    for bad_tile in (1, 3, 48, 128, -64):  # the store / load benchmarks' tile widths
        assert L.trove_store_counting(fake, 10, 20, 1, bad_tile, None) == T.ERR_VARIANT
        assert L.trove_load_sum(fake, fake, 10, 20, 1, bad_tile, None) == T.ERR_VARIANT
    assert L.trove_block_copy(fake, fake, 10, 17, 0, 0, None) == T.ERR_INVALID
    assert L.trove_block_copy(fake, fake + 4, 10, 4, 0, 0, None) == T.ERR_ALIGN
    assert L.trove_block_write(fake, 10, 0, None) == T.ERR_INVALID
    assert L.trove_memcpy_ceiling(fake, fake, 24, None) == T.ERR_INVALID
    assert L.trove_random_lines_ceiling(fake, 0, fake, fake, 10, None) == T.ERR_INVALID
    assert [T.scatter_permutation_inverts(b) for b in (4, 8, 12, 32, 60, 64, 96, 6)] == \
        [False, False, True, True, True, False, False, False]
    assert L.trove_scatter_permutation_inverts(68) == T.ERR_SIZE
    assert L.trove_invert_permutation(fake, fake, -1, None) == T.ERR_INVALID
    assert L.trove_invert_permutation(fake + 2, fake, 10, None) == T.ERR_ALIGN
    assert L.trove_invert_permutation(0, 0, 0, None) == T.OK
    assert L.trove_scatter_permutation(fake, fake, fake, 10, 68, fake, None) == T.ERR_SIZE
    assert L.trove_scatter_permutation(fake, fake, fake, 10, 20, fake + 2, None) == T.ERR_ALIGN
    assert L.trove_scatter_permutation(fake, 0, fake, 10, 20, fake, None) == T.ERR_INVALID
    assert L.trove_random_rows_write_ceiling(fake, 100, 48, fake, 10, 0, None) == T.ERR_SIZE
    assert L.trove_random_rows_write_ceiling(fake, 0, 64, fake, 10, 0, None) == T.ERR_INVALID
    assert L.trove_random_rows_write_ceiling(fake + 4, 100, 64, fake, 10, 0, None) == T.ERR_ALIGN
    assert L.trove_random_rows_write_ceiling(fake + 4, 100, 60, fake + 2, 10, 1, None) == T.ERR_ALIGN
    assert L.trove_random_lines_ceiling(fake, 1 << 32, fake, fake, 10, None) == T.ERR_INVALID
    assert L.trove_random_lines_ceiling(fake + 64, 8, fake, fake, 10, None) == T.ERR_ALIGN
    assert L.trove_random_lines_ceiling(fake, 8, fake, fake, -1, None) == T.ERR_INVALID
    assert L.trove_random_lines_ceiling(0, 8, 0, 0, 0, None) == T.OK
    with pytest.raises(T.TroveError):
        T.copy(fake, fake, 10, 68, stream=0)


def test_generators_are_deterministic():
    a = T.gen_permutation(10000, 7)
    assert np.array_equal(np.sort(a), np.arange(10000))
    assert np.array_equal(a, T.gen_permutation(10000, 7))
    assert not np.array_equal(a, T.gen_permutation(10000, 8))
    w = T.gen_words(1000, 5)
    assert np.array_equal(w, T.gen_words(1000, 5))
    assert len(np.unique(w)) > 990


def test_permutation_is_uniform_enough():
    # Position of element 0 over many seeds: roughly uniform over 8 buckets.
    n = 8
    counts = np.zeros(n)
    for seed in range(4000):
        counts[int(np.argmax(T.gen_permutation(n, seed) == 0))] += 1
    assert counts.min() > 400 and counts.max() < 600, counts


def test_device_schedule_on_host(tmp_path):
    exe = str(tmp_path / "schedule_check")
    subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "host", "schedule_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr


def test_reference_style_kernels_compile(tmp_path):
    proc = subprocess.run([HIPCC, "-std=c++17", "-O3", "--offload-arch=gfx950", "-Werror=nontrivial-memcall",
                           "-I" + os.path.join(ROOT, "include"), "-c",
                           os.path.join(ROOT, "tests", "host", "dropin_usage.hip"), "-o", str(tmp_path / "d.o")],
                          capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-4000:]


def test_host_drivers_reject_bad_arguments_before_any_gpu_call():
    host = os.path.join(ROOT, "tests", "host")
    for args in (["trove_benchmark", "--n", "0"], ["trove_benchmark", "--upper", "17"], ["trove_block", "--s", "17"]):
        exe = os.path.join(host, args[0])
        assert os.path.exists(exe), f"{exe} not built"
        p = subprocess.run([exe, *args[1:]], capture_output=True, text=True, timeout=60)
        assert p.returncode == 2, (args, p.stdout, p.stderr)
