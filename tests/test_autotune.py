"""The autotuner's knob (host-only calls, no GPU): the fused FFA kernels' tile rows are
validated and read back; the default is restored with 0."""

from __future__ import annotations

import pytest

from pyloki_amd import autotune


def test_tile_rows_round_trip():
    try:
        for rows in (1, 8, 16, 64, 256):
            autotune.set_tile_rows(rows)
            assert autotune.get_tile_rows() == rows
    finally:
        autotune.set_tile_rows(0)
    assert autotune.get_tile_rows() == 0


@pytest.mark.parametrize("rows", [3, -1, 512, 24])
def test_tile_rows_rejects_invalid(rows):
    with pytest.raises(ValueError, match="power of two"):
        autotune.set_tile_rows(rows)
    assert autotune.get_tile_rows() == 0


def test_cache_path_override(tmp_path, monkeypatch):
    monkeypatch.setenv("PYLOKI_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    assert autotune.cache_path() == tmp_path / "t.json"
    autotune._store({"dev": {"ffa_tile_rows/x": {"rows": 8, "ms": {}}}})
    assert autotune._load()["dev"]["ffa_tile_rows/x"]["rows"] == 8
