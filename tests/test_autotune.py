"""Autotuner cache logic on the CPU (a stand-in cell records what is applied)."""
import json

import numpy as np

from voxel_lmfem import autotune


class FakeTorch:
    class cuda:
        @staticmethod
        def get_device_name(device):
            return "fake-gpu"


class FakeIndex:
    def element_size(self):
        return 1


class FakeCell:
    def __init__(self, shape=(32, 32, 32)):
        self.shape = shape
        self.index = FakeIndex()
        self.scale = None
        self.device = None
        self.torch = FakeTorch
        self.config = (14, 2, 8)
        self.tunes = 0

    def stencil_config(self, precision="float32"):
        return self.config

    def set_stencil_config(self, precision, tile_y, variant, zmin):
        if tile_y not in (14, 8, 6):
            raise ValueError("unsupported tile")
        self.config = (tile_y, variant, zmin)

    def autotune(self, precision="float32", nrhs=6, reps=3):
        self.tunes += 1
        table = np.arange(1.0, 28.0).reshape(3, 3, 3)[::-1, ::-1, ::-1].copy()  # fastest: last entry
        self.config = (6, 3, 2)
        return (6, 3, 2), table


def test_tune_records_and_reuses(tmp_path):
    path = tmp_path / "t.json"
    cell = FakeCell()
    first = autotune.tune(cell, "float32", 6, path=path)
    assert first["tuned"] and cell.tunes == 1
    assert (first["tile_y"], first["schedule"], first["zmin"]) == (6, 3, 2)
    assert first["ms"] == 1.0
    # default_ms is the untuned configuration's entry: tile 14 (axis 0 index 0),
    # schedule 2, 8 planes (axis 2 index 0) -> table[0, 1, 0] of the reversed ramp
    assert first["default_ms"] == 27.0 - (0 * 9 + 1 * 3 + 0)
    data = json.loads(path.read_text())
    assert data["schema"] == autotune.SCHEMA
    assert list(data["entries"]) == ["fake-gpu|fp32|32x32x32|r6|u8|plain"]
    other = FakeCell()
    second = autotune.tune(other, "float32", 6, path=path)
    assert not second["tuned"] and other.tunes == 0 and other.config == (6, 3, 2)
    # a different grid, precision or load count is a different key
    assert autotune.tune(FakeCell((16, 32, 32)), "float32", 6, path=path)["tuned"]
    assert autotune.tune(FakeCell(), "float64", 6, path=path)["tuned"]
    assert autotune.tune(FakeCell(), "float32", 1, path=path)["tuned"]
    assert len(autotune.load(path)) == 4


def test_stale_entry_and_bad_file(tmp_path):
    path = tmp_path / "t.json"
    path.write_text("not json")
    assert autotune.load(path) == {}
    cell = FakeCell()
    key = autotune.key(cell, "float32", 6)
    autotune.save({key: dict(tile_y=16, schedule=2, zmin=8)}, path)  # tile this build does not have
    got = autotune.tune(cell, "float32", 6, path=path)
    assert got["tuned"] and cell.tunes == 1
    assert autotune.load(path)[key]["tile_y"] == 6


def test_shared_entries_are_not_written(tmp_path, monkeypatch):
    monkeypatch.setenv("VXL_TUNE_CACHE", str(tmp_path / "env.json"))
    entries = {}
    autotune.tune(FakeCell(), "float32", 6, entries=entries)
    assert len(entries) == 1 and not (tmp_path / "env.json").exists()
    autotune.save(entries)
    assert (tmp_path / "env.json").is_file()


def test_default_ms_follows_the_cells_configuration(tmp_path):
    """The recorded default time is the entry of whatever configuration the
    cell ran before tuning (schedule 3 by default in both precisions)."""
    table = np.arange(27.0).reshape(3, 3, 3)
    for precision, config, expect in (("float32", (14, 3, 8), table[0, 2, 0]),
                                      ("float64", (6, 3, 8), table[0, 2, 0]),
                                      ("float64", (4, 1, 2), table[2, 0, 2]),
                                      ("float32", (8, 2, 4), table[1, 1, 1])):
        assert autotune._table_ms(table, precision, *config) == expect
    assert np.isnan(autotune._table_ms(table, "float32", 16, 3, 8))
