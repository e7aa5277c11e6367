"""Tile-table handling of the autotuner on the CPU: nearest-shape lookup per precision,
and the load cache following the file on disk."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from pynof_amd import autotune  # noqa: E402


def test_lookup_nearest_shape_same_precision():
    table = {"split:100:360:100": {"half_transform": [2, 2]}, "split:1000:3600:300": {"half_transform": [5, 5]},
             "fp64:1000:3600:300": {"half_transform": [5, 4]}}
    assert autotune.lookup(table, "split", 900, 3240, 280) == {"half_transform": [5, 5]}
    assert autotune.lookup(table, "split", 120, 432, 120) == {"half_transform": [2, 2]}
    assert autotune.lookup(table, "fp64", 120, 432, 120) == {"half_transform": [5, 4]}
    assert autotune.lookup(table, "other", 120, 432, 120) is None


def test_load_follows_saves(tmp_path):
    p = str(tmp_path / "tiles.json")
    assert autotune.load(p) == {}
    autotune.save({"split:1:8:1": {"exchange_kn": [2, 2]}}, p)
    first = autotune.load(p)
    first["added"] = {}  # the caller's copy; the cached table is unchanged
    assert autotune.load(p) == {"split:1:8:1": {"exchange_kn": [2, 2]}}
    autotune.save({"fp64:2:8:2": {"exchange_kn": [4, 4]}}, p)
    assert autotune.load(p) == {"fp64:2:8:2": {"exchange_kn": [4, 4]}}


def test_committed_table_parses():
    table = autotune.load()
    assert table and all(k.split(":")[0] in ("fp64", "split") for k in table)
    for entry in table.values():
        assert set(entry) <= set(autotune.STAGES)
        for st, (fm, fn) in entry.items():
            assert (fm, fn) in autotune.CANDIDATES[st]
