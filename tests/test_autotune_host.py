"""Autotuner cache on the host (no GPU): versioned JSON, atomic store,
keys that separate architectures, geometries and flags."""

import json

from nlse_amd import autotune


def test_cache_roundtrip(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("NLSE_AMD_TUNE_CACHE", str(path))
    assert autotune.load() == {}
    autotune.store("plan|gfx950|a", {"col_width": 8, "graph_chunk": 16})
    autotune.store("plan|gfx950|b", {"col_width": 4, "graph_chunk": 0})
    assert autotune.lookup("plan|gfx950|a")["col_width"] == 8
    assert autotune.lookup("plan|gfx950|b")["graph_chunk"] == 0
    data = json.loads(path.read_text())
    assert data["version"] == autotune.VERSION and len(data["entries"]) == 2
    # another version of the cache format is ignored, not misread
    path.write_text(json.dumps({"version": autotune.VERSION + 1, "entries": {"x": {}}}))
    assert autotune.load() == {}
    path.write_text("not json")
    assert autotune.load() == {}


def test_keys_distinguish_shapes_and_flags(monkeypatch):
    monkeypatch.setattr(autotune, "arch", lambda: "gfx950")
    geo = dict(ndim=2, ny=2048, nx=2048, nt=1, items=1, ncomp=1)
    k0 = autotune.plan_key(geo, False, None)
    assert k0 != autotune.plan_key(geo, True, None)
    assert k0 != autotune.plan_key(geo, False, (25, 25))
    assert k0 != autotune.plan_key(dict(geo, ny=1024), False, None)
    assert k0 != autotune.plan_key(dict(geo, ncomp=2), False, None)
    assert autotune.batch_key([(256, 256), (512, 512)]) == autotune.batch_key([(512, 512), (256, 256)])
    assert autotune.batch_key([(256, 256)]) != autotune.batch_key([(256, 256), (256, 256)])


def test_enabled_flag(monkeypatch):
    monkeypatch.delenv("NLSE_AMD_AUTOTUNE", raising=False)
    assert not autotune.enabled()
    monkeypatch.setenv("NLSE_AMD_AUTOTUNE", "1")
    assert autotune.enabled()
