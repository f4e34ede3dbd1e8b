"""Persistence of autotuned choices (CPU only): save/load round trip per
device and class, the user's file before the shipped one."""
import json

from argent_amd import autotune


def test_save_load_round_trip(tmp_path):
    path = str(tmp_path / "t.json")
    choice = {"tuning": {"paths_per_wave": 1 << 24, "blocks_per_cu": 4}, "ms_start": 10.0, "ms_best": 9.0,
              "trials": 12}
    autotune.save("preview", choice, path)
    autotune.save("still", {"tuning": {"blocks_per_cu": 5}, "ms_start": 1, "ms_best": 1, "trials": 1}, path)
    t, src = autotune.load("preview", path)
    assert t == choice["tuning"] and src.startswith("user:")
    assert json.load(open(path))["gfx950"]["still"]["tuning"] == {"blocks_per_cu": 5}
    t, src = autotune.load("no-such-class", path)
    assert t is None or src.startswith("shipped:")


def test_shipped_choices_are_valid_tunings():
    import os
    if not os.path.exists(autotune.SHIPPED):
        return
    data = json.load(open(autotune.SHIPPED))
    for cls, entry in data["gfx950"].items():
        assert set(entry["tuning"]) <= set(autotune.FIELDS), cls
        assert entry["ms_best"] <= entry["ms_start"], cls
