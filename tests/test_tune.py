"""Launch tunables and the autotuner's cache (host only: no kernel runs)."""
import json

import pytest

from neuralangelo_amd import _lib, autotune


def test_names_defaults_and_ranges():
    lib = _lib.load()
    names = [lib.na_tune_name(i).decode() for i in range(lib.na_tune_count())]
    assert set(names) == set(autotune.CANDIDATES)
    assert lib.na_tune_name(lib.na_tune_count()) is None
    for name in names:
        old = autotune.get(name)
        try:
            for v in autotune.CANDIDATES[name]:
                autotune.set(name, v)
                assert autotune.get(name) == v
        finally:
            autotune.set(name, old)
    with pytest.raises(RuntimeError, match="out of range"):
        autotune.set("taps_bwd.run", 0)
    with pytest.raises(RuntimeError, match="multiple"):
        autotune.set("packed.rays_per_wg", 96)
    with pytest.raises(RuntimeError, match="unknown tunable"):
        autotune.get("no.such.knob")


def test_cache_is_applied_and_validated(tmp_path):
    lib = _lib.load()
    old = autotune.get("taps_bwd.run")
    path = tmp_path / "cache.json"
    try:
        path.write_text(json.dumps({"gfx950": {"taps_bwd.run": {"value": 7}}, "gfx942": {
            "taps_bwd.run": {"value": 3}}}))
        autotune.apply_cache(lib, str(path))
        assert autotune.get("taps_bwd.run") == 7  # only this build's architecture
        path.write_text(json.dumps({"gfx950": {"taps_bwd.run": 99}}))
        with pytest.raises(RuntimeError, match="rejected"):
            autotune.apply_cache(lib, str(path))
    finally:
        autotune.set("taps_bwd.run", old)


def test_committed_cache_covers_every_tunable():
    cache = autotune.load_cache()
    entries = cache.get(autotune.ARCH, {})
    for name, rec in entries.items():
        assert name in autotune.CANDIDATES
        assert rec["value"] in autotune.CANDIDATES[name]
        assert str(rec["value"]) in rec["timings_ms"]
        assert min(rec["timings_ms"].values()) == rec["timings_ms"][str(rec["value"])]
