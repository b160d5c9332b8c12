"""Tile-choice persistence of the autotuner (CPU only)."""
import importlib


def test_cache_roundtrip(tmp_path, monkeypatch):
    monkeypatch.setenv("CGS_TUNING_FILE", str(tmp_path / "t.json"))
    from cugridsearch_amd import tuner
    importlib.reload(tuner)
    assert tuner.load() == {}
    tuner.save({"plain:64x64:u8:cost1:sampler1": 8})
    assert tuner.load()["plain:64x64:u8:cost1:sampler1"] == 8
    importlib.reload(tuner)       # a fresh process reads the file
    assert tuner.load()["plain:64x64:u8:cost1:sampler1"] == 8
