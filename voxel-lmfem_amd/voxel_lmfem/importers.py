"""Import an external labelled voxel volume into the solve input format.

Same contract as the reference's ``import_model``
(src/voxel_lmfem/importers.py:66-157, README.md:124-133): the output is a new
directory holding ``model.npz`` (``phase_id`` in z, y, x order,
``domain_size`` in x, y, z order, and the orientation arrays),
``materials.json`` (the validated table) and ``import.json`` (schema 1: source
paths with SHA-256, source and normalised axes and shapes, physical size,
spacing, voxel count and volume fraction per phase).  Geometry is preserved
exactly, only permuted; vectors and tensors stay in global x, y, z
components.  A rejected model leaves no output directory behind.
"""
import hashlib
import json
import tempfile
from pathlib import Path

import numpy as np

from . import materials as mat
from .inputs import validate

_REQUIRED = {"isotropic": {"E", "nu"},
             "orthotropic": {"E1", "E2", "E3", "nu12", "nu13", "nu23", "G12", "G13", "G23"},
             "anisotropic": {"C"}}
_I64 = np.iinfo(np.int64)


def sha256_file(path, block=1 << 20):
    h = hashlib.sha256()
    with Path(path).open("rb") as f:
        while True:
            chunk = f.read(block)
            if not chunk:
                return h.hexdigest()
            h.update(chunk)


def _directions(archive):
    """orientation | axis1+axis2 | None from an open NPZ."""
    names = set(archive.files)
    pair = names & {"axis1", "axis2"}
    if "orientation" in names:
        if pair:
            raise ValueError("Use orientation or axis1/axis2, rather than both representations")
        return archive["orientation"]
    if pair:
        if pair != {"axis1", "axis2"}:
            raise ValueError("Both axis1 and axis2 are required")
        return {name: archive[name] for name in ("axis1", "axis2")}
    return None


def _load_volume(path, key):
    suffix = path.suffix.lower()
    if suffix == ".npy":
        return np.load(path, allow_pickle=False), None, None
    if suffix == ".npz":
        with np.load(path, allow_pickle=False) as archive:
            if key not in archive.files:
                raise ValueError(f"NPZ has no {key!r} array; select its label array with --key")
            domain = archive["domain_size"] if "domain_size" in archive.files else None
            return archive[key], domain, _directions(archive)
    raise ValueError("Supported volume formats: .npy and .npz")


def _load_directions(path):
    suffix = path.suffix.lower()
    if suffix == ".npy":
        return np.load(path, allow_pickle=False)
    if suffix == ".npz":
        with np.load(path, allow_pickle=False) as archive:
            found = _directions(archive)
        if found is None:
            raise ValueError("Orientation NPZ needs orientation or axis1/axis2 arrays")
        return found
    raise ValueError("Orientation input must be .npy or .npz")


def _permute_directions(directions, source_shape, order):
    """Spatial axes of a direction field follow the label array; components do not."""
    if directions is None:
        return None

    def one(values, tail):
        a = np.asarray(values, dtype=float)
        if a.shape == tail:
            return a
        if a.shape != tuple(source_shape) + tail:
            raise ValueError(f"Orientation shape must be {tail} or {tuple(source_shape) + tail} in source axis order")
        return a.transpose(tuple(order) + tuple(range(3, a.ndim)))

    if isinstance(directions, dict):
        return {name: one(directions[name], (3,)) for name in ("axis1", "axis2")}
    return one(directions, (3, 3))


def _check_materials(table):
    if not isinstance(table, dict) or not table or not all(isinstance(v, dict) for v in table.values()):
        raise ValueError("Materials JSON must map phase IDs to material property objects")
    for label, spec in table.items():
        kind = spec.get("kind")
        if kind not in _REQUIRED:
            raise ValueError(f"Material {label}: kind must be isotropic, orthotropic or anisotropic")
        missing = sorted(_REQUIRED[kind] - set(spec))
        if missing:
            raise ValueError(f"Material {label}: missing {', '.join(missing)}")
        if not _I64.min <= int(label) <= _I64.max:
            raise ValueError(f"Material ID {label} must fit a signed 64-bit integer")


def import_model(input_path, materials_path, out, *, axes="zyx", spacing=None, domain_size=None, key="phase_id",
                 orientation_path=None):
    """Convert and validate an external labelled volume; returns the import report."""
    source, materials_path, target = Path(input_path), Path(materials_path), Path(out).resolve()
    if target.exists():
        raise FileExistsError(f"Choose a new output directory: {target}")
    if not isinstance(axes, str) or sorted(axes) != ["x", "y", "z"]:
        raise ValueError("axes must be a permutation of xyz, such as zyx or xyz")
    if spacing is not None and domain_size is not None:
        raise ValueError("Specify spacing or domain_size, rather than both")
    labels, stored_domain, directions = _load_volume(source, key)
    if labels.ndim != 3 or labels.dtype.kind not in "biu" or min(labels.shape) < 2:
        raise ValueError("Input must be a 3D integer label array with every dimension >= 2; "
                         "segment image intensities first")
    if labels.dtype.kind == "b":
        labels = labels.astype(np.uint8)
    if labels.dtype.kind == "u" and labels.size and int(labels.max()) > _I64.max:
        raise ValueError("Material IDs must fit a signed 64-bit integer")
    if orientation_path is not None:
        if directions is not None:
            raise ValueError("Orientation is already present in the source NPZ")
        orientation_path = Path(orientation_path)
        directions = _load_directions(orientation_path)
    source_shape = labels.shape
    order = [axes.index(a) for a in "zyx"]   # source axis feeding z, y, x
    phase = np.ascontiguousarray(labels.transpose(order)).astype(labels.dtype.newbyteorder("="), copy=False)
    directions = _permute_directions(directions, source_shape, order)
    if spacing is not None:
        spacing = np.asarray(spacing, dtype=float)
        if spacing.shape != (3,) or not np.isfinite(spacing).all() or (spacing <= 0).any():
            raise ValueError("Spacing must be three finite positive values in x,y,z order")
        domain_size = spacing * np.array(phase.shape[::-1], dtype=float)
    elif domain_size is None:
        domain_size = stored_domain
    if domain_size is None:
        raise ValueError("Provide --spacing or --domain-size, or domain_size in the source NPZ")
    table = json.loads(materials_path.read_text(encoding="utf-8-sig"))
    _check_materials(table)
    try:
        phase, checked, domain, directions, labels_sorted, _, local = validate(phase, table, domain_size, directions)
    except (TypeError, KeyError, OverflowError) as exc:
        raise ValueError(f"Invalid material definition: {exc}") from exc
    flat = phase.reshape(-1)
    for k in np.flatnonzero(local):   # every local frame must be a proper rotation
        where = np.flatnonzero(flat == labels_sorted[k])
        for first in range(0, len(where), 65536):
            mat.frames(directions, phase.shape, where[first:first + 65536])
    arrays = dict(phase_id=phase, domain_size=domain)
    if isinstance(directions, dict):
        arrays.update(directions)
    elif directions is not None:
        arrays["orientation"] = directions
    ids, counts = np.unique(phase, return_counts=True)
    sources = {"volume": source, "materials": materials_path}
    if orientation_path is not None:
        sources["orientation"] = orientation_path
    report = dict(schema_version=1, input_axes=axes, input_shape=list(source_shape), axes="zyx",
                  shape=list(phase.shape), domain_size=domain.tolist(),
                  spacing=(domain / np.array(phase.shape[::-1], dtype=float)).tolist(),
                  phases=[dict(id=int(i), voxels=int(c), fraction=float(c / phase.size)) for i, c in zip(ids, counts)],
                  sources={name: dict(file=str(p.resolve()), sha256=sha256_file(p)) for name, p in sources.items()})
    table_text = json.dumps({str(k): v for k, v in checked.items()}, indent=2, allow_nan=False) + "\n"
    target.parent.mkdir(parents=True, exist_ok=True)
    with tempfile.TemporaryDirectory(prefix=".vxl-import-", dir=target.parent) as staging:
        stage = Path(staging)
        np.savez_compressed(stage / "model.npz", **arrays)
        (stage / "materials.json").write_text(table_text, encoding="utf-8")
        (stage / "import.json").write_text(json.dumps(report, indent=2, allow_nan=False) + "\n", encoding="utf-8")
        if target.exists():
            raise FileExistsError(f"Choose a new output directory: {target}")
        stage.rename(target)
    return report
