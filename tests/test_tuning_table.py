"""The committed per-grid tiling table (inq_amd/tuning.json "hpsi") and its
lookup, on the CPU."""
import pytest

from inq_amd import tuning


def test_band_classes_cover_every_band_count():
    assert tuning.band_class(1) == "1-32"
    assert tuning.band_class(32) == "1-32"
    assert tuning.band_class(33) == "33-128"
    assert tuning.band_class(128) == "33-128"
    assert tuning.band_class(129) == "129+"
    assert tuning.band_class(4096) == "129+"
    with pytest.raises(ValueError):
        tuning.band_class(0)


def test_chunk_candidates_are_pipeline_widths():
    for L, nst in ((64, 256), (96, 512), (24, 16), (48, 64)):
        c = tuning.chunk_candidates(L, nst)
        assert c[0] == 0
        assert all(b % 8 == 0 and 2 * b <= nst for b in c[1:])


def test_committed_table_is_consistent():
    table = tuning.load()
    assert table, "tuning.json has no per-grid table"
    assert tuning.check_table_consistency(table)
    # the grid lengths of the workloads are covered in every band class
    for L in (24, 30, 32, 36, 40, 48, 60, 64, 72, 96):
        assert str(L) in table, L
        assert set(table[str(L)]) == {"1-32", "33-128", "129+"}, L


def test_lookup(tmp_path):
    table = tuning.load()
    assert tuning.lookup((64, 64, 48), 256) is None          # not cubic
    assert tuning.lookup((17, 17, 17), 16) is None           # no entry
    k = tuning.lookup((64, 64, 64), 256)
    assert set(k) == set(tuning.KNOBS)
    assert k["band_chunk"] == table["64"]["129+"]["band_chunk"]
    # a chunk wider than the handle's bands falls back to the library's choice
    import json
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"hpsi": {"48": {"33-128": {
        "z_tile_y": 16, "xy_threads": 1, "band_chunk": 64, "projector_bands": 32}}}}))
    assert tuning.lookup((48, 48, 48), 40, path=str(path)) == {
        "z_tile_y": 16, "xy_threads": 1, "band_chunk": 0, "projector_bands": 32}
    assert tuning.lookup((48, 48, 48), 100, path=str(path))["band_chunk"] == 64
