"""The import-model contract (reference tests/test_import.py): normalised
model.npz / materials.json / import.json, axis permutation of labels and
direction fields, SHA-256 provenance, and rejected inputs leaving nothing."""
import json
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from voxel_lmfem.importers import import_model, sha256_file
from voxel_lmfem.inputs import load_input


def materials_file(tmp_path, table):
    path = tmp_path / "materials.json"
    path.write_text(json.dumps(table), encoding="utf-8")
    return path


def cli(*args, cwd):
    env_path = str(ROOT / "voxel-lmfem_amd")
    return subprocess.run([sys.executable, "-c", "import sys; sys.path.insert(0, sys.argv[1]); "
                           "from voxel_lmfem.cli import main; sys.exit(main(sys.argv[2:]))", env_path,
                           *[str(a) for a in args]], cwd=cwd, capture_output=True, text=True)


def test_npy_cli_import(tmp_path):
    source = tmp_path / "external.npy"
    np.save(source, np.full((2, 3, 4), 17, dtype=">u2"))
    materials = materials_file(tmp_path, {17: dict(kind="isotropic", E=3500., nu=.35)})
    target = tmp_path / "converted"
    result = cli("import-model", "--input", source, "--materials", materials, "--axes", "xyz",
                 "--spacing", .5, 2, 3, "--out", target, cwd=tmp_path)
    assert result.returncode == 0, result.stderr
    assert json.loads(result.stdout)["status"] == "imported"
    phase, table, domain, _ = load_input(target / "model.npz", target / "materials.json")
    assert phase.shape == (4, 3, 2) and phase.dtype.isnative
    np.testing.assert_array_equal(domain, [1., 6., 12.])
    report = json.loads((target / "import.json").read_text())
    assert report["schema_version"] == 1 and report["axes"] == "zyx" and report["input_axes"] == "xyz"
    assert report["phases"] == [dict(id=17, voxels=24, fraction=1.)]
    assert report["sources"]["volume"]["sha256"] == sha256_file(source)
    assert len(report["sources"]["materials"]["sha256"]) == 64
    before = (target / "model.npz").read_bytes()
    with pytest.raises(FileExistsError):
        import_model(source, materials, target, spacing=[1, 1, 1])
    assert (target / "model.npz").read_bytes() == before
    again = cli("import-model", "--input", source, "--materials", materials, "--spacing", 1, 1, 1,
                "--out", target, cwd=tmp_path)
    assert again.returncode == 2 and again.stderr.startswith("error:")


def test_spatial_labels_and_directions_follow_axis_permutation(tmp_path):
    phase = np.zeros((2, 3, 4), dtype=np.uint8)
    phase[1, 2, 3] = 1
    rotation = np.broadcast_to(np.eye(3), (2, 3, 4, 3, 3)).copy()
    turned = np.array([[0., -1., 0.], [1., 0., 0.], [0., 0., 1.]])
    rotation[1, 2, 3] = turned
    source = tmp_path / "external.npz"
    np.savez(source, labels=phase, orientation=rotation, domain_size=[2, 3, 4])
    materials = materials_file(tmp_path, {0: dict(kind="isotropic", E=1., nu=.2),
                                         1: dict(kind="anisotropic", C=np.diag([10, 20, 30, 4, 5, 6]).tolist(),
                                                 frame="local")})
    target = tmp_path / "converted"
    import_model(source, materials, target, axes="xyz", key="labels")
    actual, _, domain, directions = load_input(target / "model.npz", target / "materials.json")
    assert actual.shape == (4, 3, 2) and actual[3, 2, 1] == 1 and actual.sum() == 1
    np.testing.assert_array_equal(directions[3, 2, 1], turned)
    np.testing.assert_array_equal(domain, [2, 3, 4])


def test_external_axis_pair_file(tmp_path):
    source, directions = tmp_path / "phase.npy", tmp_path / "axes.npz"
    np.save(source, np.zeros((2, 2, 2), dtype=np.uint8))
    np.savez(directions, axis1=[0, 1, 0], axis2=[-1, 0, 0])
    materials = materials_file(tmp_path, {0: dict(kind="anisotropic", frame="local",
                                                C=np.diag([10., 20., 30., 4., 5., 6.]).tolist())})
    target = tmp_path / "converted"
    report = import_model(source, materials, target, domain_size=[1, 2, 3], orientation_path=directions)
    assert set(report["sources"]) == {"volume", "materials", "orientation"}
    _, _, domain, orientation = load_input(target / "model.npz", target / "materials.json")
    np.testing.assert_array_equal(orientation["axis1"], [0, 1, 0])
    np.testing.assert_array_equal(domain, [1, 2, 3])


@pytest.mark.parametrize("case,message", [
    ("float", "integer label"), ("missing-size", "spacing"), ("unknown-phase", "Undefined material"),
    ("bad-size", "positive"), ("missing-direction", "require orientation"),
    ("bad-frame", "orthonormal"), ("ambiguous-direction", "rather than both"),
])
def test_rejected_models_leave_no_partial_output(tmp_path, case, message):
    source = tmp_path / "external.npz"
    phase = np.zeros((2, 2, 2), dtype=float if case == "float" else np.uint8)
    if case == "unknown-phase":
        phase[:] = 2
    arrays = dict(phase_id=phase)
    if case != "missing-size":
        arrays["domain_size"] = [1, 0 if case == "bad-size" else 1, 1]
    material = dict(kind="isotropic", E=1., nu=.2)
    if case in ("missing-direction", "bad-frame", "ambiguous-direction"):
        material = dict(kind="anisotropic", C=np.eye(6).tolist(), frame="local")
    if case == "bad-frame":
        arrays["orientation"] = np.ones((3, 3))
    if case == "ambiguous-direction":
        arrays.update(orientation=np.eye(3), axis1=[1, 0, 0], axis2=[0, 1, 0])
    np.savez(source, **arrays)
    materials = materials_file(tmp_path, {0: material})
    with pytest.raises(ValueError, match=message):
        import_model(source, materials, tmp_path / "converted")
    assert not (tmp_path / "converted").exists()
    assert not list(tmp_path.glob(".vxl-import-*"))
