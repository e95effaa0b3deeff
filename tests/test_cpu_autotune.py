"""Autotuner bookkeeping (host only): candidates, workload keys, cache file."""
import json

from deepdrr_amd import autotune
from deepdrr_amd._abi_types import View


def _view(w, h):
    v = View()
    v.width, v.height, v.n_volumes = w, h, 1
    return v


def test_candidates_are_valid_and_distinct():
    assert all(autotune.valid(c) for c in autotune.CANDIDATES)
    assert len(set(autotune.CANDIDATES)) == len(autotune.CANDIDATES)
    assert not autotune.valid((8, 4, 4, 8)) and not autotune.valid((8, 8, 5, 8))
    assert not autotune.valid((8, 8, 4, 0)) and not autotune.valid((8, 8, 4, 33))


def test_workload_key_ignores_poses_not_sizes():
    a = autotune.workload_key([_view(64, 64)] * 5, [0] * 5, [(10, 11, 12)], 3, "gfx950")
    b = autotune.workload_key([_view(64, 64)] * 6, [0] * 6, [(10, 11, 12)], 3, "gfx950")
    c = autotune.workload_key([_view(64, 32)] * 5, [0] * 5, [(10, 11, 12)], 3, "gfx950")
    d = autotune.workload_key([_view(64, 64)] * 5, [0] * 5, [(10, 11, 12)], 5, "gfx950")
    e = autotune.workload_key([_view(64, 64)] * 5, [0] * 5, [(10, 11, 12)], 3, "gfx950", march="fast")
    assert a == b  # same size bucket (5 and 6 -> 8)
    assert a != c and a != d and a != e
    assert json.loads(a)["det"] == [[64, 64]]


def test_cache_file_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"k1": [16, 4, 2, 8], "bad": [8, 4, 4, 8]}))
    monkeypatch.setenv("DEEPDRR_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(autotune, "_cache", {})
    monkeypatch.setattr(autotune, "_file_loaded", False)
    assert autotune.cached("k1") == (16, 4, 2, 8)
    assert autotune.cached("bad") is None
    with autotune._lock:
        autotune._cache["k2"] = (8, 8, 1, 16)
        autotune._save_file()
    assert json.loads(path.read_text())["k2"] == [8, 8, 1, 16]
