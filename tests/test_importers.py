"""import-model: the contract of the reference's importer (messages, the
import.json schema 1 fields, SHA-256 provenance, no partial output), checked
on volumes built here.

Axis normalisation is checked point by point: a source voxel at source index
(i0, i1, i2) with source axes e.g. "xzy" is the voxel x=i0, z=i1, y=i2, and
must appear at phase_id[z, y, x]; direction fields must travel with it while
their components stay global.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from voxel_lmfem.importers import file_sha256, import_model
from voxel_lmfem.inputs import load_input

ISO = dict(kind="isotropic", E=3500.0, nu=0.35)
TI_LOCAL = dict(kind="orthotropic", frame="local", E1=230.0, E2=15.0, E3=15.0, nu12=0.2, nu13=0.2,
                nu23=0.3, G12=27.0, G13=27.0, G23=15.0 / 2.6)


def write_json(path, value):
    path.write_text(json.dumps(value), encoding="utf-8")
    return path


def leftovers(parent):
    """Staging directories an import may have left beside its target."""
    return [p.name for p in parent.iterdir() if p.name.startswith(".vxl-import-")]


def normalised_index(axes, source_index):
    """zyx index of a source voxel, from the axis names alone."""
    where = dict(zip(axes, source_index))
    return where["z"], where["y"], where["x"]


def random_rotations(rng, count):
    q = rng.normal(size=(count, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


@pytest.mark.parametrize("axes", ["xyz", "xzy", "yxz", "yzx", "zxy", "zyx"])
def test_labels_follow_axis_names(tmp_path, axes):
    rng = np.random.default_rng(len(axes) + sum(map(ord, axes)))
    src_shape = (3, 4, 5)
    labels = rng.integers(0, 3, size=src_shape).astype(np.int16)
    source = tmp_path / "labels.npy"
    np.save(source, labels)
    mats = write_json(tmp_path / "m.json", {str(k): ISO for k in range(3)})
    spacing = [0.5, 0.25, 2.0]
    report = import_model(source, mats, tmp_path / "out", axes=axes, spacing=spacing)
    phase, table, domain, orientation = load_input(tmp_path / "out" / "model.npz", tmp_path / "out" / "materials.json")
    for idx in np.ndindex(*src_shape):
        assert phase[normalised_index(axes, idx)] == labels[idx]
    nz, ny, nx = phase.shape
    assert {axes[k]: src_shape[k] for k in range(3)} == {"z": nz, "y": ny, "x": nx}
    np.testing.assert_array_equal(domain, np.array(spacing) * [nx, ny, nz])
    assert orientation is None and set(table) == {"0", "1", "2"}
    assert report["input_axes"] == axes and report["input_shape"] == list(src_shape)
    assert report["axes"] == "zyx" and report["shape"] == [nz, ny, nx]
    np.testing.assert_allclose(report["spacing"], spacing)


def test_report_schema_and_provenance(tmp_path):
    phase = np.zeros((2, 4, 5), dtype=">u2")
    phase[:, 1:3, :] = 7
    phase[1, 0, 0] = 9
    source = tmp_path / "vol.npz"
    np.savez(source, phase_id=phase, domain_size=np.array([5.0, 4.0, 2.0]))
    mats = write_json(tmp_path / "m.json", {"0": ISO, "7": ISO, "9": ISO})
    target = tmp_path / "nested" / "model"
    report = import_model(source, mats, target)
    on_disk = json.loads((target / "import.json").read_text())
    assert on_disk == report
    assert set(report) == {"schema_version", "input_axes", "input_shape", "axes", "shape", "domain_size",
                           "spacing", "phases", "sources"}
    assert report["schema_version"] == 1
    assert report["domain_size"] == [5.0, 4.0, 2.0] and report["spacing"] == [1.0, 1.0, 1.0]
    census = {p["id"]: (p["voxels"], p["fraction"]) for p in report["phases"]}
    assert census == {0: (19, 19 / 40), 7: (20, 0.5), 9: (1, 1 / 40)}
    assert [p["id"] for p in report["phases"]] == [0, 7, 9]
    for role, path in (("volume", source), ("materials", mats)):
        entry = report["sources"][role]
        assert entry["file"] == str(path.resolve())
        assert entry["sha256"] == hashlib.sha256(path.read_bytes()).hexdigest()
    model = np.load(target / "model.npz", allow_pickle=False)
    assert model["phase_id"].dtype == np.dtype("uint16") and model["phase_id"].dtype.isnative
    np.testing.assert_array_equal(model["phase_id"], phase)
    assert sorted(os.listdir(target)) == ["import.json", "materials.json", "model.npz"]
    assert leftovers(target.parent) == []


def test_file_sha256_chunking(tmp_path):
    data = np.random.default_rng(3).bytes(3 * 1000 + 17)
    path = tmp_path / "blob"
    path.write_bytes(data)
    for chunk in (1, 7, 1000, 1 << 20):
        assert file_sha256(path, chunk) == hashlib.sha256(data).hexdigest()
    path.write_bytes(b"")
    assert file_sha256(path) == hashlib.sha256(b"").hexdigest()


def test_boolean_labels_become_uint8(tmp_path):
    mask = np.zeros((2, 3, 2), dtype=bool)
    mask[0, 1, 1] = True
    source = tmp_path / "mask.npy"
    np.save(source, mask)
    mats = write_json(tmp_path / "m.json", {"0": ISO, "1": ISO})
    import_model(source, mats, tmp_path / "out", domain_size=[1, 1, 1])
    phase = np.load(tmp_path / "out" / "model.npz")["phase_id"]
    assert phase.dtype == np.uint8 and phase.sum() == 1


def test_rotation_field_travels_with_voxels(tmp_path):
    rng = np.random.default_rng(5)
    axes, src_shape = "yzx", (3, 2, 4)
    labels = np.ones(src_shape, dtype=np.int32)
    labels[0] = 2
    rot = random_rotations(rng, np.prod(src_shape)).reshape(src_shape + (3, 3))
    source = tmp_path / "vol.npy"
    np.save(source, labels)
    rot_path = tmp_path / "rot.npy"
    np.save(rot_path, rot)
    mats = write_json(tmp_path / "m.json", {"1": TI_LOCAL, "2": ISO})
    report = import_model(source, mats, tmp_path / "out", axes=axes, domain_size=[1, 2, 3],
                          orientation_path=rot_path)
    phase, _, _, orientation = load_input(tmp_path / "out" / "model.npz", tmp_path / "out" / "materials.json")
    for idx in np.ndindex(*src_shape):
        z, y, x = normalised_index(axes, idx)
        assert phase[z, y, x] == labels[idx]
        np.testing.assert_array_equal(orientation[z, y, x], rot[idx])   # components untouched
    assert report["sources"]["orientation"]["sha256"] == file_sha256(rot_path)


def test_axis_pair_in_source_npz(tmp_path):
    axes, src_shape = "xzy", (4, 2, 3)
    rng = np.random.default_rng(9)
    labels = np.full(src_shape, 4, dtype=np.uint8)
    a1 = rng.normal(size=src_shape + (3,))
    a2 = np.broadcast_to(np.array([0.0, 0.0, 1.0]), (3,))   # one vector for every voxel stays (3,)
    a1[..., 2] = 0.0   # never collinear with axis2
    source = tmp_path / "vol.npz"
    np.savez(source, phase_id=labels, axis1=a1, axis2=a2)
    mats = write_json(tmp_path / "m.json", {"4": TI_LOCAL})
    import_model(source, mats, tmp_path / "out", axes=axes, spacing=[1, 1, 1])
    model = np.load(tmp_path / "out" / "model.npz")
    assert model["axis2"].shape == (3,)
    for idx in np.ndindex(*src_shape):
        np.testing.assert_array_equal(model["axis1"][normalised_index(axes, idx)], a1[idx])


def test_uniform_rotation_is_kept_as_one_matrix(tmp_path):
    source = tmp_path / "vol.npz"
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    np.savez(source, phase_id=np.zeros((2, 2, 3), dtype=np.int64), orientation=R)
    mats = write_json(tmp_path / "m.json", {"0": TI_LOCAL})
    import_model(source, mats, tmp_path / "out", axes="zxy", spacing=[1, 1, 1])
    np.testing.assert_array_equal(np.load(tmp_path / "out" / "model.npz")["orientation"], R)


def _npz(path, **arrays):
    np.savez(path, **arrays)
    return path


REJECTED = [
    # (case, kwargs-builder, exception, message fragment)
    ("axes", lambda d: dict(axes="xxz", spacing=[1, 1, 1]), ValueError, "axes must be a permutation of xyz"),
    ("axes-length", lambda d: dict(axes="xyzx", spacing=[1, 1, 1]), ValueError, "axes must be a permutation"),
    ("spacing-and-size", lambda d: dict(spacing=[1, 1, 1], domain_size=[1, 1, 1]), ValueError,
     "Specify spacing or domain_size, rather than both"),
    ("format", lambda d: dict(input_path=d / "vol.tif", spacing=[1, 1, 1]), ValueError,
     "Supported volume formats: .npy and .npz"),
    ("key", lambda d: dict(input_path=_npz(d / "k.npz", labels=np.zeros((2, 2, 2), np.uint8)), spacing=[1, 1, 1]),
     ValueError, "NPZ has no 'phase_id' array; select its label array with --key"),
    ("two-representations", lambda d: dict(input_path=_npz(d / "both.npz", phase_id=np.zeros((2, 2, 2), np.uint8),
                                                            orientation=np.eye(3), axis1=np.eye(3)[0]),
                                           spacing=[1, 1, 1]),
     ValueError, "Use orientation or axis1/axis2, rather than both representations"),
    ("half-pair", lambda d: dict(input_path=_npz(d / "half.npz", phase_id=np.zeros((2, 2, 2), np.uint8),
                                                 axis2=np.eye(3)[0]), spacing=[1, 1, 1]),
     ValueError, "Both axis1 and axis2 are required"),
    ("float-volume", lambda d: dict(input_path=_npz(d / "f.npz", phase_id=np.zeros((2, 2, 2))), spacing=[1, 1, 1]),
     ValueError, "Input must be a 3D integer label array"),
    ("thin-volume", lambda d: dict(input_path=_npz(d / "t.npz", phase_id=np.zeros((1, 2, 2), np.uint8)),
                                   spacing=[1, 1, 1]), ValueError, "every dimension >= 2"),
    ("huge-id", lambda d: dict(input_path=_npz(d / "u.npz", phase_id=np.full((2, 2, 2), 2 ** 63, np.uint64)),
                               spacing=[1, 1, 1]), ValueError, "Material IDs must fit a signed 64-bit integer"),
    ("orientation-twice", lambda d: dict(input_path=_npz(d / "o.npz", phase_id=np.zeros((2, 2, 2), np.uint8),
                                                         orientation=np.eye(3)),
                                         orientation_path=d / "extra.npy", spacing=[1, 1, 1]),
     ValueError, "Orientation is already present in the source NPZ"),
    ("orientation-format", lambda d: dict(orientation_path=d / "rot.txt", spacing=[1, 1, 1]), ValueError,
     "Orientation input must be .npy or .npz"),
    ("orientation-npz-empty", lambda d: dict(orientation_path=_npz(d / "r.npz", other=np.eye(3)), spacing=[1, 1, 1]),
     ValueError, "Orientation NPZ needs orientation or axis1/axis2 arrays"),
    ("orientation-shape", lambda d: dict(orientation_path=_save(d / "r.npy", np.zeros((2, 2, 3, 3, 3))),
                                         spacing=[1, 1, 1]),
     ValueError, "Orientation shape must be (3, 3) or (2, 2, 2, 3, 3) in source axis order"),
    ("spacing-bad", lambda d: dict(spacing=[1, 0, 1]), ValueError,
     "Spacing must be three finite positive values in x,y,z order"),
    ("no-size", lambda d: dict(), ValueError, "Provide --spacing or --domain-size, or domain_size in the source NPZ"),
    ("domain-bad", lambda d: dict(domain_size=[1, -1, 1]), ValueError, "domain_size must contain three finite"),
    ("materials-list", lambda d: dict(materials_path=write_json(d / "l.json", [ISO]), spacing=[1, 1, 1]),
     ValueError, "Materials JSON must map phase IDs to material property objects"),
    ("materials-kind", lambda d: dict(materials_path=write_json(d / "k.json", {"0": dict(kind="cubic")}),
                                      spacing=[1, 1, 1]),
     ValueError, "Material 0: kind must be isotropic, orthotropic or anisotropic"),
    ("materials-missing", lambda d: dict(materials_path=write_json(d / "mm.json", {"0": dict(kind="orthotropic",
                                                                                              E1=1.0)}),
                                         spacing=[1, 1, 1]),
     ValueError, "Material 0: missing E2, E3, G12, G13, G23, nu12, nu13, nu23"),
    ("materials-id", lambda d: dict(materials_path=write_json(d / "id.json", {str(2 ** 63): ISO}), spacing=[1, 1, 1]),
     ValueError, "must fit a signed 64-bit integer"),
    ("materials-undefined", lambda d: dict(materials_path=write_json(d / "u.json", {"1": ISO}), spacing=[1, 1, 1]),
     ValueError, "Undefined material labels: [0]"),
    ("materials-nu", lambda d: dict(materials_path=write_json(d / "nu.json", {"0": dict(kind="isotropic", E=1.0,
                                                                                         nu=0.5)}),
                                    spacing=[1, 1, 1]),
     ValueError, "Isotropic elasticity requires E>0 and -1<nu<0.5"),
    ("local-without-frames", lambda d: dict(materials_path=write_json(d / "loc.json", {"0": TI_LOCAL}),
                                            spacing=[1, 1, 1]),
     ValueError, "Local-frame anisotropic phases require orientation"),
    ("improper-rotation", lambda d: dict(materials_path=write_json(d / "loc2.json", {"0": TI_LOCAL}),
                                         orientation_path=_save(d / "ref.npy", np.diag([1.0, 1.0, -1.0])),
                                         spacing=[1, 1, 1]),
     ValueError, "Orientation must be right-handed (det=+1)"),
]


def _save(path, array):
    np.save(path, array)
    return path


@pytest.mark.parametrize("case,build,exc,message", REJECTED, ids=[r[0] for r in REJECTED])
def test_rejected_inputs_leave_nothing(tmp_path, case, build, exc, message):
    np.save(tmp_path / "vol.npy", np.zeros((2, 2, 2), dtype=np.uint8))
    np.save(tmp_path / "extra.npy", np.eye(3))
    args = dict(input_path=tmp_path / "vol.npy", materials_path=write_json(tmp_path / "ok.json", {"0": ISO}))
    args.update(build(tmp_path))
    target = tmp_path / "out"
    before = sorted(os.listdir(tmp_path))
    with pytest.raises(exc) as err:
        import_model(args.pop("input_path"), args.pop("materials_path"), target, **args)
    assert message in str(err.value)
    assert not target.exists()
    assert sorted(os.listdir(tmp_path)) == before and leftovers(tmp_path) == []


def test_existing_target_is_refused_and_untouched(tmp_path):
    np.save(tmp_path / "vol.npy", np.zeros((2, 2, 2), dtype=np.uint8))
    mats = write_json(tmp_path / "m.json", {"0": ISO})
    target = tmp_path / "out"
    import_model(tmp_path / "vol.npy", mats, target, spacing=[1, 1, 1])
    snapshot = {name: (target / name).read_bytes() for name in os.listdir(target)}
    np.save(tmp_path / "vol.npy", np.ones((2, 2, 2), dtype=np.uint8))
    with pytest.raises(FileExistsError, match="Choose a new output directory"):
        import_model(tmp_path / "vol.npy", write_json(tmp_path / "m1.json", {"1": ISO}), target, spacing=[2, 2, 2])
    assert {name: (target / name).read_bytes() for name in os.listdir(target)} == snapshot
    assert leftovers(tmp_path) == []


def test_target_appearing_during_import_is_not_replaced(tmp_path, monkeypatch):
    """The existence check is repeated just before the rename: a directory
    created meanwhile is kept and the staged files are removed."""
    from voxel_lmfem import importers
    np.save(tmp_path / "vol.npy", np.zeros((2, 2, 2), dtype=np.uint8))
    mats = write_json(tmp_path / "m.json", {"0": ISO})
    target = tmp_path / "out"
    original = importers._json_writer

    def racing_writer(value):
        write = original(value)

        def go(path):
            write(path)
            target.mkdir(exist_ok=True)
        return go
    monkeypatch.setattr(importers, "_json_writer", racing_writer)
    with pytest.raises(FileExistsError):
        import_model(tmp_path / "vol.npy", mats, target, spacing=[1, 1, 1])
    assert target.is_dir() and os.listdir(target) == []
    assert leftovers(tmp_path) == []
