"""Host logic of the launch-shape autotuner: cache keys, the JSON cache, the choice."""
import json

from iterative_tomographic_reconstruction_algorithms_amd import autotune as T


def test_bucket_key_groups_nearby_geometries():
    k = T.bucket_key("MI355X", "dd", "fast", 2048, 1800, 1)
    assert k == "MI355X|dd|fast|n2048|v2048|z1"
    assert T.bucket_key("MI355X", "dd", "fast", 2000, 1025, 1) == k
    assert T.bucket_key("MI355X", "dd", "fast", 2049, 1800, 1) != k
    assert T.bucket_key("MI355X", "dd", "ieee", 2048, 1800, 1) != k
    assert T.bucket_key("MI355X", "dd", "fast", 2048, 1800, 3) == "MI355X|dd|fast|n2048|v2048|z4"


def test_cache_round_trip_and_atomic_overwrite(tmp_path):
    p = str(tmp_path / "sub" / "shapes.json")
    assert T.load_cache(p) == {}
    T.store_entry("a", {"fwd": 1, "back": 2}, p)
    T.store_entry("b", {"fwd": 0, "back": 0}, p)
    T.store_entry("a", {"fwd": 2, "back": 1}, p)
    data = json.load(open(p))
    assert data == {"a": {"fwd": 2, "back": 1}, "b": {"fwd": 0, "back": 0}}
    assert [f.name for f in (tmp_path / "sub").iterdir()] == ["shapes.json"]


def test_corrupt_cache_reads_as_empty(tmp_path):
    p = tmp_path / "shapes.json"
    p.write_text("{not json")
    assert T.load_cache(str(p)) == {}
    p.write_text("[1, 2]")
    assert T.load_cache(str(p)) == {}


def test_cache_path_from_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("ITOMO_SHAPE_CACHE", str(tmp_path / "x.json"))
    assert T.cache_path() == str(tmp_path / "x.json")
    assert T.cache_path("/explicit.json") == "/explicit.json"


def test_fastest_takes_first_minimum():
    assert T.fastest([3.0, 1.0, 1.0]) == 1
    assert T.fastest([0.5]) == 0
    assert T.fastest([2.0, 2.0, 3.0]) == 0
