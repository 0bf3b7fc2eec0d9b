"""Tile table (no GPU): keys by size class, the library's rule without an entry, and
entries read back from the saved table."""
import json

from packing_amd import autotune as A


def test_size_class_keys():
    assert A.size_class(10) == "R1" and A.size_class(200) == "R10" and A.size_class(212) == "R10"
    assert A.size_class(213) == "N256" and A.size_class(300) == "N512"
    assert A.key("lut", 200, 10, 8) == "lut:R10:P10:m8"
    assert list(A.candidates(200, 8)) == [1, 2, 64, 128, 256]
    assert list(A.candidates(213, 8)) == [64, 128, 256]
    assert list(A.candidates(100, 12)) == [64, 128, 256]


def test_choose_block_reads_table(tmp_path, monkeypatch):
    path = tmp_path / "tiles.json"
    path.write_text(json.dumps({"lut:R10:P10:m8": {"block": 128, "ms": {"1": 2.0, "128": 1.0}},
                                "area:R2:P4:m8": 64}))
    monkeypatch.setattr(A, "TABLE_PATH", str(path))
    monkeypatch.setattr(A, "_table", None)
    assert A.choose_block("lut", 200, 10, 8) == 128
    assert A.choose_block("area", 30, 4, 8) == 64
    assert A.choose_block("lut", 50, 10, 8) == 0          # no entry: the library's rule
    monkeypatch.setattr(A, "_table", None)


def test_bucket_count_is_limited():
    """At most four concurrent size buckets (one stream per hardware queue)."""
    import pytest
    from packing_amd.relax import RelaxOptions
    RelaxOptions(buckets=4)
    RelaxOptions(buckets=1)
    with pytest.raises(ValueError):
        RelaxOptions(buckets=6)
