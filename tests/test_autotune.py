"""Autotuner host logic: candidate selection and the on-disk cache (no GPU)."""
import json

import pytest

from pyatmosphere_amd import autotune


@pytest.fixture
def cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("PA_TUNE_CACHE", str(path))
    # a fixed device identity, so no HIP device is touched
    monkeypatch.setattr(autotune, "_prefix", lambda device: "testdev|lib1")
    return path


def test_pick_takes_the_fastest_and_the_first_on_ties():
    assert autotune.pick({1: 2.0, 2: 1.5, 3: 1.7}) == 2
    assert autotune.pick({16: 1.0, 8: 1.0, 4: 3.0}) == 16
    with pytest.raises(ValueError):
        autotune.pick({})


def test_cache_roundtrip_and_validation(cache):
    assert autotune.lookup(2048, 1024, 32, "f16x3", None) is None
    key = autotune.screen_key(2048, 1024, 32, "f16x3", None)
    autotune._store(key, {"gemm_tile": 1, "table_rows": 16})
    assert autotune.lookup(2048, 1024, 32, "f16x3", None)["gemm_tile"] == 1
    # other shapes and modes miss
    assert autotune.lookup(2048, 1024, 16, "f16x3", None) is None
    assert autotune.lookup(2048, 1024, 32, "bf16x3", None) is None
    # entries naming shapes the library does not have are ignored
    autotune._store(autotune.screen_key(1024, 1024, 8, "f16x3", None), {"gemm_tile": 3, "table_rows": 16})
    assert autotune.lookup(1024, 1024, 8, "f16x3", None) is None
    autotune._store(autotune.batch_key(2048, 1024, 5, "f16x3", None), {"batch": 32, "overlap": False})
    assert autotune.lookup_batch(2048, 1024, 5, "f16x3", None) == {"batch": 32, "overlap": False}
    assert autotune.lookup_batch(2048, 1024, 4, "f16x3", None) is None
    d = json.loads(cache.read_text())
    assert len(d) == 3


def test_corrupt_cache_reads_as_empty(cache):
    cache.write_text("{not json")
    assert autotune.lookup(512, 1024, 8, "f16x3", None) is None
    autotune._store(autotune.screen_key(512, 1024, 8, "f16x3", None), {"gemm_tile": 2, "table_rows": 8})
    assert autotune.lookup(512, 1024, 8, "f16x3", None)["table_rows"] == 8
