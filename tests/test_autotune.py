"""The autotuner's JSON cache (no device): records round-trip by key, a
missing or corrupt file reads as empty, and a cached choice is returned
with every field."""
import json

import pytest

from tool_godot_curvenet_amd import autotune as A


class _Info:
    n_free, nnz_ff = 1000, 6800


class _Sys:
    info = _Info()


@pytest.fixture
def fake_device(monkeypatch):
    monkeypatch.setattr(A, "_device", lambda: "testdev/gfx950/256cu")


def test_cache_round_trip(tmp_path, fake_device):
    path = str(tmp_path / "t.json")
    key = A.key_of(_Sys(), 768, "mixed", "hsc")
    assert "n=1000" in key and "k=768" in key and "testdev" in key
    assert A.lookup(_Sys(), 768, "mixed", "hsc", path) is None
    A._store(path, key, {"tune": {"spmm_form": 2, "level_cap": 4096}, "ms": 40.0})
    got = A.lookup(_Sys(), 768, "mixed", "hsc", path)
    assert got == {"spmm_form": 2, "spmm_gens": 0, "level_cap": 4096, "tail_rows": 0,
                   "chain_levels": 0, "coarse_prec": 0, "down_cap": 0}
    # another k, precision or preconditioner is another key
    assert A.lookup(_Sys(), 48, "mixed", "hsc", path) is None
    assert A.lookup(_Sys(), 768, "fp64", "hsc", path) is None
    # a second record keeps the first
    A._store(path, A.key_of(_Sys(), 48, "mixed", "hsc"), {"tune": {"tail_rows": -1}})
    assert len(json.load(open(path))) == 2


def test_corrupt_cache_reads_empty(tmp_path, fake_device):
    path = tmp_path / "bad.json"
    path.write_text("{not json")
    assert A.lookup(_Sys(), 3, path=str(path)) is None
    path.write_text("[1, 2]")
    assert A._load(str(path)) == {}


def test_cache_path_env(monkeypatch, tmp_path):
    monkeypatch.setenv("CN_AUTOTUNE_CACHE", str(tmp_path / "x.json"))
    assert A.cache_path() == str(tmp_path / "x.json")


def test_tail_applies():
    assert A._tail_applies(768, "mixed") and A._tail_applies(512, "mixed")
    assert not A._tail_applies(48, "mixed") and not A._tail_applies(3, "mixed")
    assert A._tail_applies(256, "fp64") and not A._tail_applies(254, "fp64")
