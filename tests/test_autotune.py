"""Autotuner host logic (no GPU): cache keys and the cache file."""
import json

from g2s_amd import autotune


def test_geometry_key():
    assert autotune.geometry_key((500, 500), [0, 1], 60) == "D2:500x500:t01:n60"
    assert autotune.geometry_key((180, 150, 120), [1], 100) == "D3:180x150x120:t1:n100"
    assert autotune.geometry_key((250,), [0], 12) != autotune.geometry_key((250,), [1], 12)


def test_cache_roundtrip(tmp_path):
    path = str(tmp_path / "tune.json")
    assert autotune.load_cache(path) == {}
    data = {"D2:90x70:t0:n20": {"scan.grid1024_min": 1025, "fft.radix_max": 8, "fft.lognv_max": 6,
                                "pencil.npb": 2}}
    autotune.save_cache(path, data)
    assert autotune.load_cache(path) == data
    with open(path) as fh:
        assert json.load(fh) == data


def test_defaults_cover_every_candidate_key():
    keys = {k for k, _ in autotune.OWN_GRID_CANDIDATES} | {"scan.grid1024_min", "loop.graph"}
    assert keys == set(autotune.DEFAULTS)
    for k, values in autotune.OWN_GRID_CANDIDATES:
        assert autotune.DEFAULTS[k] in values
