"""CPU checks of the batched-gather tuner's table logic (trove_amd/tune.py)."""
import json

from trove_amd import tune


def test_bucket_is_nearest_on_log_scale():
    assert tune.bucket(1) == 1 << 10
    assert tune.bucket(1 << 12) == 1 << 10 or tune.bucket(1 << 12) == 1 << 14
    assert tune.bucket((1 << 14) + 5) == 1 << 14
    assert tune.bucket(1 << 17) == 1 << 18
    assert tune.bucket(1 << 30) == 1 << 22


def test_choose_reads_the_table_and_falls_back():
    table = {"version": 1, "lengths": [1024, 1 << 20],
             "entries": [{"bytes": 20, "pattern": "random", "length": 1024, "variant": "shfl", "chunk_groups": 1},
                         {"bytes": 20, "pattern": "random", "length": 1 << 20, "variant": "tile",
                          "chunk_groups": 64}]}
    assert tune.choose(table, 20, "random", 900) == ("shfl", 1)
    assert tune.choose(table, 20, "random", 1 << 19) == ("tile", 64)
    # No entry: tile path for multi-word structs, per-lane for one-word ones.
    assert tune.choose(table, 36, "runs", 5000) == ("tile", 16)
    assert tune.choose(table, 16, "runs", 5000) == ("tile", 16)


def test_table_round_trip(tmp_path):
    table = {"version": 1, "lengths": list(tune.LENGTHS), "entries": []}
    p = str(tmp_path / "t.json")
    tune.save(table, p)
    assert tune.load(p) == table
    assert json.load(open(p))["version"] == 1


def test_variant_candidates_per_size():
    # One-word structs have no transposition: per-lane loads (tile policies
    # or naive), or the LDS stage for dwordx4 stores; every other size tries
    # all four paths.
    assert tune._variants(4) == ("tile", "lds", "direct")
    assert tune._variants(16) == ("tile", "lds", "direct")
    assert set(tune._variants(20)) == {"tile", "shfl", "lds", "direct"}
