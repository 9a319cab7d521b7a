"""The autotuner's shipped table and how it is applied (no GPU)."""
import json
import os

import mlsmpm
from mlsmpm import autotune


def test_shipped_table_is_well_formed():
    raw = json.load(open(autotune.TABLE))
    for key in ("batch_ppt_by_class", "alone_ppt_by_class"):
        t = autotune.load_table(key=key)
        assert t and all(isinstance(k, int) for k in t)
        assert set(t.values()) <= {4, 8, 16}
        assert set(t) == {int(k) for k in raw[key]}
    # the classes cover the benchmark mix: 10k .. 2M particles
    t = autotune.load_table()
    assert autotune.size_class(10_000) in t and autotune.size_class(1_999_999) in t


def test_size_class_is_two_octaves():
    assert autotune.size_class(1) == 0
    assert autotune.size_class(4) == 1 and autotune.size_class(15) == 1
    assert autotune.size_class(16) == 2
    assert autotune.size_class(1 << 20) == 10 and autotune.size_class((1 << 22) - 1) == 10


def test_apply_tiles_sets_only_unset_sizes_from_the_class():
    table = {"7": 4, 9: 16}
    c = autotune.apply_tiles(mlsmpm.default_config(64, 64, 64, 40_000), table)
    assert c.tile_ppt == 4                      # 40k -> class 7
    c = autotune.apply_tiles(mlsmpm.default_config(64, 64, 64, 300_000), table)
    assert c.tile_ppt == 16                     # 300k -> class 9
    c = autotune.apply_tiles(mlsmpm.default_config(64, 64, 64, 150_000), table)
    assert c.tile_ppt == 0                      # class 8: no entry, default
    c = autotune.apply_tiles(mlsmpm.default_config(64, 64, 64, 40_000, tile_ppt=16), table)
    assert c.tile_ppt == 16                     # a size the caller named stays
    c = autotune.apply_tiles(mlsmpm.default_config(64, 64, 64, 40_000), None)
    assert c.tile_ppt == 0


def test_missing_table_reads_as_none(tmp_path):
    assert autotune.load_table(os.path.join(tmp_path, "none.json")) is None
    p = os.path.join(tmp_path, "t.json")
    autotune.save(p, {"batch_ppt_by_class": {"3": 8}})
    assert autotune.load_table(p) == {3: 8}
    assert autotune.load_table(p, key="alone_ppt_by_class") is None
